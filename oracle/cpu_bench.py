"""CPU baseline of the RK4 step: the reference's CPU path restated in oracle/phx_cpu_rk4.c, timed on
the host cores.  TEST INFRASTRUCTURE ONLY (bench.py's cpu_baseline leg runs it as a child process).

Protocol: for each precision and each subgrid split {4x4, 8x8, 16x16, 32x32} of the 1024^2 workload
(scalar GP + reservoir, periodic, ring pump 1.5 P_th, dt 0.04, psi0 = 0.01 complex Gaussian noise),
20 warm-up steps, then the median of 200 individually timed steps; the best split is reported.  The
256^2 FP64 workload (Gaussian pump, 1000 steps) is timed as well.  The OpenMP thread binding is the
reference benchmark's (OMP_PROC_BIND=true, OMP_PLACES=cores), which the caller sets in the child's
environment before the OpenMP runtime starts.

    python3 oracle/cpu_bench.py [--size 1024] [--steps 200] [--warmup 20]  ->  one JSON line
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PTH = 0.15 * 0.225 / 0.01


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def isa_level() -> str:
    """x86-64-v4 where the host has the AVX-512 subsets of that level, else v3."""
    try:
        flags = set()
        for line in open("/proc/cpuinfo"):
            if line.startswith("flags"):
                flags = set(line.split(":", 1)[1].split())
                break
        if {"avx512f", "avx512bw", "avx512vl", "avx512dq", "avx512cd"} <= flags:
            return "v4"
    except OSError:
        pass
    return "v3"


def load(prec: str, level: str):
    L = C.CDLL(os.path.join(HERE, f"libphx_cpu_rk4_{prec}_{level}.so"))
    dp = C.POINTER(C.c_double)
    L.cpu_rk4_setup.argtypes = [C.c_int, C.c_int, dp, dp, dp] + [C.c_double] * 11
    L.cpu_rk4_run.argtypes = [C.c_int, dp]
    L.cpu_rk4_get.argtypes = [dp, dp]
    return L


def constants(N: int, L: float):
    """SystemParameters::calculateAuto (system.cpp:116-160) with the default physics, as the oracle
    derives them (orc_derive)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("phx_oracle_wrapper", os.path.join(HERE, "oracle.py"))
    O = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(O)
    p = O.params(N, N, L, L, pump=True)
    return {k: getattr(p, k) for k in ("m_eff_scaled", "one_over_h_bar_s", "m2_over_dx2_p_dy2", "one_over_dx2",
                                       "one_over_dy2", "g_c", "g_r", "R", "gamma_c", "gamma_r")}


def workload(N: int, ring: bool, seed: int):
    L = 0.5 * (N - 1)
    x = -L / 2 + 0.5 * np.arange(N)
    X, Y = np.meshgrid(x, x)
    r = np.hypot(X, Y)
    P = 1.5 * PTH * (np.exp(-((r - 40.0) ** 2) / 50.0) if ring else np.exp(-r ** 2 / 200.0))
    rng = np.random.default_rng(seed)
    psi = 0.01 * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
    return L, np.ascontiguousarray(psi), np.zeros((N, N)), np.ascontiguousarray(P)


def setup(lib, N, split, psi, n, P, dt, k):
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    rc = lib.cpu_rk4_setup(N, split, dp(psi.view(np.float64)), dp(n), dp(P), dt, k["m_eff_scaled"], k["one_over_h_bar_s"],
                           k["m2_over_dx2_p_dy2"], k["one_over_dx2"], k["one_over_dy2"], k["g_c"], k["g_r"], k["R"],
                           k["gamma_c"], k["gamma_r"])
    if rc:
        raise RuntimeError(f"cpu_rk4_setup({N}, {split}) failed: {rc}")


def time_steps(lib, steps: int, warmup: int) -> float:
    if warmup:
        lib.cpu_rk4_run(warmup, None)
    us = np.zeros(steps)
    lib.cpu_rk4_run(steps, us.ctypes.data_as(C.POINTER(C.c_double)))
    return float(np.median(us))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--splits", default="4,8,16,32")
    ap.add_argument("--small-steps", type=int, default=1000)
    a = ap.parse_args()
    level = isa_level()
    N = a.size
    L, psi, n, P = workload(N, True, 4242)
    k = constants(N, L)
    out = {"cpu": cpu_model(), "isa": f"x86-64-{level}", "threads": None,
           "omp": {v: os.environ.get(v, "") for v in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}}
    for prec in ("fp32", "fp64"):
        lib = load(prec, level)
        out["threads"] = lib.cpu_rk4_threads()
        res = {}
        for split in [int(s) for s in a.splits.split(",") if s]:
            setup(lib, N, split, psi, n, P, 0.04, k)
            res[f"{split}x{split}"] = round(time_steps(lib, a.steps, a.warmup), 1)
        lib.cpu_rk4_free()
        best = min(res, key=res.get)
        out[prec] = {"median_us_per_step": res, "best_split": best, "best_us": res[best],
                     "gcell_steps_per_s": round(N * N / (res[best] * 1e-6) / 1e9, 4)}
    # the 256^2 FP64 workload, 1000 steps: one subgrid (the reference's GPU auto rule for 256^2)
    lib = load("fp64", level)
    Ls, psis, ns, Ps = workload(256, False, 1234)
    setup(lib, 256, 1, psis, ns, Ps, 0.04, constants(256, Ls))
    out["fp64_256_1000_steps_ms"] = round(1e-3 * time_steps(lib, a.small_steps, 0) * a.small_steps, 1)
    lib.cpu_rk4_free()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    sys.exit(main())
