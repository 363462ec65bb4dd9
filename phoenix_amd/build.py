"""Build libphx.so and the PHOENIX-compatible executables for gfx950.

Usage: python -m phoenix_amd.build [-i] [-v]   (or phoenix_amd.build.build() from Python)
  -i  recompile only objects older than their source or a header

Outputs (in-tree, so they travel with the source tree):
  phoenix_amd/lib/libphx.so            C ABI of include/phx.h
  phoenix_amd/bin/PHOENIX_gpu_fp32     CLI, same flags as the reference executables
  phoenix_amd/bin/PHOENIX_gpu_fp64
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")

ARCH = os.environ.get("PHX_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# Packed FP32 VALU ops stay disabled on every device pass.  No automatic FMA contraction:
# the kernels write their fused multiply-adds explicitly (phx_fma), so results are
# bitwise independent of the launch shape.
HIP_FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wno-unused-result", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", os.path.join(ROOT, "include")]
# The streaming kernels keep their row windows in registers only when the row loop is
# unrolled completely; the 7-stage spinor kernels exceed LLVM's default pragma threshold, and so
# does the FP64 RK4 kernel at 24 rows with 8 prefetched rows (4 KiB of scratch per lane without it).
_UNROLL = ["-mllvm", "-pragma-unroll-threshold=2000000"]
HIP_FLAGS_FILE = {f: _UNROLL for f in ("phx_rk_generic.hip", "phx_rk_tab3.hip", "phx_rk_tab4.hip", "phx_rk_tab6.hip",
                                     "phx_rk_tab6f.hip", "phx_rk_dp45t.hip", "phx_rk4_f64.hip")}
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-I", os.path.join(ROOT, "include")]

# longest compiles first: the thread pool then finishes them side by side
HIP_SOURCES = ["phx_rk_generic.hip", "phx_rk_dp45t.hip", "phx_rk4_f64.hip", "phx_rk4.hip", "phx_rk4_flow.hip", "phx_rk4_flow_f64.hip", "phx_rk_tab6.hip", "phx_rk_tab6f.hip", "phx_rk_tab4.hip",
               "phx_rk_tab3.hip", "phx_fft.hip", "phx_solver.hip", "phx_ensemble.hip", "phx_adjoint.hip", "phx_ssfm.hip",
               "phx_staged.hip"]
CXX_SOURCES = ["phx_config.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def _includes(path: str, seen: set) -> set:
    """Files reached from `path` through quoted #include lines (the library's own headers)."""
    import re
    if path in seen or not os.path.exists(path):
        return seen
    seen.add(path)
    with open(path) as f:
        for m in re.finditer(r'^\s*#\s*include\s+"([^"]+)"', f.read(), re.M):
            for base in (os.path.dirname(path), os.path.join(ROOT, "include")):
                cand = os.path.normpath(os.path.join(base, m.group(1)))
                if os.path.exists(cand):
                    _includes(cand, seen)
                    break
    return seen


def _stale(obj: str, src: str) -> bool:
    """Object missing or older than its source, this script or a header the source includes."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = _includes(src, set()) | {__file__}
    return any(os.path.getmtime(d) > t for d in deps)


def build(verbose: bool = False, incremental: bool = False) -> str:
    """Compile every source (incremental=True: only objects older than their inputs)."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    jobs = []
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s.replace(".hip", ".o"))
        if not incremental or _stale(o, os.path.join(CSRC, s)):
            jobs.append([HIPCC, *HIP_FLAGS, *HIP_FLAGS_FILE.get(s, []), "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in CXX_SOURCES:
        o = os.path.join(BUILD, s.replace(".cpp", ".o"))
        if not incremental or _stale(o, os.path.join(CSRC, s)):
            jobs.append([CXX, *CXX_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    if not jobs:
        jobs.append(["true"])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for r in ex.map(_run, jobs):
            if verbose and r.stderr:
                print(r.stderr, file=sys.stderr)
    lib = os.path.join(LIB, "libphx.so")
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, lib)
    main_o = os.path.join(BUILD, "phx_main.o")
    _run([CXX, *CXX_FLAGS, "-c", os.path.join(CSRC, "phx_main.cpp"), "-o", main_o])
    for prec in ("fp32", "fp64"):
        exe = os.path.join(BIN, f"PHOENIX_gpu_{prec}")
        _run([CXX, f"-DPHX_MAIN_PRECISION={0 if prec == 'fp32' else 1}", *CXX_FLAGS,
              os.path.join(CSRC, "phx_main.cpp"), "-o", exe, f"-L{LIB}", "-lphx", f"-Wl,-rpath,$ORIGIN/../lib"])
    return lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="-i" in sys.argv))
