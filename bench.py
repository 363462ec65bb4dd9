#!/usr/bin/env python3
"""Benchmark of the PHOENIX RK4 time step on one MI355X.

Workload (synthetic data, shapes of the 1024^2 workload): scalar GP + reservoir,
1024 x 1024 cells, dx = 0.5 um (L = 511.5 um), periodic boundaries, dt = 0.04 ps,
ring pump P = 1.5 P_th exp(-(r-40)^2 / (2*5^2)), psi0 = 0.01 complex Gaussian
noise (seed 4242), n0 = 0.  One "step" = one RK4 step over the whole grid.
The k-space filter of that workload (every 50 steps) is timed separately when
available and is not part of the per-step figure.

Prints ONE JSON line (see the contract in the repository documentation).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# Kernel arguments in device memory: the fused kernels read ~200 bytes of arguments at
# wave start, and host-resident argument buffers add about 2 us to every 10 us launch.
# Must be set before the HIP runtime initialises (before torch is imported).
os.environ.setdefault("HIP_FORCE_DEV_KERNARG", "1")
PMC_SUMMARY = os.path.join(ROOT, "profiles", "rk4_scalar_fused_1024_pmc.md")
FLOW_PMC_SUMMARY = os.path.join(ROOT, "profiles", "rk4_scalar_flow_1024_pmc.md")

METRIC = "µs/RK4 step at 1024² fp32/fp64 + % HBM roofline; batched & adjoint vs plain"
PTH = 0.15 * 0.225 / 0.01
HBM_PEAK_GBS = 8000.0


def workload_inputs(N: int, L: float, seed: int = 4242):
    dx = L / (N - 1)
    x = -L / 2 + dx * np.arange(N)
    X, Y = np.meshgrid(x, x)
    r = np.sqrt(X ** 2 + Y ** 2)
    P = 1.5 * PTH * np.exp(-((r - 40.0) ** 2) / (2 * 5.0 ** 2))
    rng = np.random.default_rng(seed)
    psi = 0.01 * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
    n = np.zeros((N, N))
    return psi, n, P


def solver_for(prec: str, N: int, L: float):
    from phoenix_amd._lib import Solver
    args = ["--N", N, N, "--L", L, L, "--boundary", "periodic", "periodic", "--tstep", 0.04, "--tmax", 1e9,
            "--fftEvery", 1e12, "--outEvery", 1e12,
            "--pump", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide"]
    s = Solver([str(a) for a in args], prec)
    psi, n, P = workload_inputs(N, L)
    s.set("psi_plus", psi)
    s.set("n_plus", n)
    s.set("pump_plus", P)
    return s


def time_rk4(prec: str, N: int, L: float, steps: int, warmup: int, kernel_samples: int = 200):
    import torch
    s = solver_for(prec, N, L)
    stream = torch.cuda.Stream()
    s.set_stream(stream.cuda_stream)
    s.step(warmup)
    s.sync()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    i0 = s.info
    t0 = time.perf_counter()
    e0.record(stream)
    done = s.step(steps)
    e1.record(stream)
    s.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total_ms = e0.elapsed_time(e1)
    i1 = s.info
    multi_steps = i1["multi_step_steps"] - i0["multi_step_steps"]
    multi_launches = i1["multi_step_launches"] - i0["multi_step_launches"]
    assert done == steps
    # per-launch duration of the fused kernel: one event pair around each single-kernel step
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(kernel_samples)]
    for a, b in evs:
        a.record(stream)
        s.step(1)
        b.record(stream)
    s.sync()
    torch.cuda.synchronize()
    k_ms = [a.elapsed_time(b) for a, b in evs] or [total_ms / steps]
    psi = s.get("psi_plus")
    finite = bool(np.isfinite(psi).all())
    info = s.info
    s.close()
    return dict(total_ms=total_ms, wall_s=wall, us_per_step=1e3 * total_ms / steps,
                kernel_us_mean=1e3 * float(np.mean(k_ms)), kernel_us_median=1e3 * float(np.median(k_ms)),
                finite=finite, tile_rows=info["tile_rows"], prefetch_rows=info["prefetch_rows"],
                multi_step_steps=multi_steps, multi_step_launches=multi_launches)


def measured_traffic(prec: str, ty: int, pf: int, kernel: str = "fused"):
    """HBM bytes per launch of the RK4 kernel for this tile shape, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summaries (lines 'traffic <prec> ty=<TY> pf=<PF>: ... bytes=<B>' for the
    one-step kernel, 'traffic flow <prec> ty=<TY> pf=<PF>: ... bytes=<B>' per step of the multi-step
    kernel, which the caller scales to one launch), or None."""
    import re
    path = PMC_SUMMARY if kernel == "fused" else FLOW_PMC_SUMMARY
    try:
        txt = open(path).read()
    except OSError:
        return None
    tag = "traffic" if kernel == "fused" else "traffic flow"
    m = re.search(rf"^{tag} {prec} ty={ty} pf={pf}:.*bytes=(\d+)", txt, re.M)
    return int(m.group(1)) if m else None


def time_fft_filter(prec: str, N: int, L: float, calls: int = 50, every: int = 50):
    """k-space low-pass filter of the 1024^2 workload: radial super-Gaussian mask (exponent 6) with
    its 1/e radius at 0.8 k_max, applied every `every` RK4 steps.  One call = forward 2D FFT, mask/N,
    inverse 2D FFT (three kernels: rows, fused columns+mask, rows)."""
    import torch
    from phoenix_amd._lib import Solver
    args = ["--N", N, N, "--L", L, L, "--boundary", "periodic", "periodic", "--tstep", 0.04, "--tmax", 1e9,
            "--fftEvery", 1e12, "--outEvery", 1e12,
            "--fftMask", 1.0, "add", 0.565685, 0.565685, 0, 0, "both", 6, "none", "gauss+noDivide+local"]
    s = Solver([str(a) for a in args], prec)
    psi, _, _ = workload_inputs(N, L)
    s.set("psi_plus", psi)
    k = (np.arange(N) - N // 2) / (N // 2)  # k / k_max, centred layout (the filter applies fftshift)
    KX, KY = np.meshgrid(k, k)
    s.set("fft_mask_plus", np.exp(-((KX ** 2 + KY ** 2) / 0.8 ** 2) ** 6))
    stream = torch.cuda.Stream()
    s.set_stream(stream.cuda_stream)
    for _ in range(5):
        s.apply_fft_filter()
    s.sync()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(calls):
        s.apply_fft_filter()
    e1.record(stream)
    s.sync()
    torch.cuda.synchronize()
    us = 1e3 * e0.elapsed_time(e1) / calls
    finite = bool(np.isfinite(s.get("psi_plus")).all())
    s.close()
    rs = 4 if prec == "fp32" else 8
    # psi is 2*rs bytes: rows (read+write) 4 rs, columns with the mask (read, mask, write) 5 rs, rows 4 rs
    moved = N * N * 13 * rs
    return {"us_per_call": round(us, 3), "every_steps": every, "us_per_step_amortized": round(us / every, 4),
            "bytes_per_call": moved, "hbm_frac": round(moved / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4), "finite": finite}


def time_spinor(N: int = 2048, steps: int = 40, prec: str = "fp32"):
    """TE-TM spinor model with adaptive Dormand-Prince 5(4), tolerance 1e-6, additive noise (--dw 1) and an
    elliptically polarised Gaussian pump (A+- = 1.5 P_th (1 +- sin 2chi) / 2, chi = pi/8), 2048^2, FP32: accepted
    steps per second.  One accepted step = one or more fused 7-stage launches, each followed by the FP64 error
    reduction and the host step-size decision."""
    import math
    import torch
    from phoenix_amd._lib import Solver
    L = 0.5 * (N - 1)
    ap = 1.5 * PTH * (1 + math.sin(math.pi / 4)) / 2
    am = 1.5 * PTH * (1 - math.sin(math.pi / 4)) / 2
    args = ["-tetm", "--N", N, N, "--L", L, L, "--boundary", "periodic", "periodic", "--iterator", "DP45", "-adaptive",
            "--tol", 1e-6, "--rkvdt", 1e-4, 0.1, "--tstep", 0.04, "--dw", 1, "--tmax", 20, "--outEvery", 1e9,
            "--fftEvery", 1e9, "--initRandom", 0.01, 7,
            "--pump", ap, "add", 10, 10, 0, 0, "plus", 1, "none", "gauss+noDivide",
            "--pump", am, "add", 10, 10, 0, 0, "minus", 1, "none", "gauss+noDivide"]
    s = Solver([str(a) for a in args], prec)
    stream = torch.cuda.Stream()
    s.set_stream(stream.cuda_stream)
    s.step(5)
    s.sync()
    i0 = s.info
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    done = s.step(steps)
    e1.record(stream)
    s.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1)
    i1 = s.info
    finite = bool(np.isfinite(s.get("psi_plus")).all() and np.isfinite(s.get("psi_minus")).all())
    s.close()
    attempts = done + (i1["rejected"] - i0["rejected"])
    us_att = 1e3 * ms / max(attempts, 1)
    bytes_att = (56 if prec == "fp32" else 112) * N * N  # psi+-, n+-, P+- read; psi+-, n+- written
    return {"grid": f"{N}x{N}", "precision": prec, "accepted_steps": done, "attempts": attempts,
            "us_per_accepted_step": round(1e3 * ms / max(done, 1), 2), "us_per_attempt": round(us_att, 2),
            "dt_end": i1["dt"], "t_end": i1["t"], "hbm_frac_per_attempt": round(bytes_att / (us_att * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
            "wall_ms": round(1e3 * wall, 2), "finite": finite}


def time_large(sizes, steps: int = 100):
    """The same workload on grids whose state no longer fits the 256 MB Infinity Cache (4096^2: 470 MB per
    FP32 step, 8192^2: 1.88 GB): µs per RK4 step and the fraction of the HBM roofline, FP32 and FP64.
    The state is too large for the multi-step kernel; steps run as replays of a 50-launch hipGraph, and
    the 60 warm-up steps (the tile autotune, then one whole block) capture it before the timed region."""
    out = {}
    for N in sizes:
        L = 0.5 * (N - 1)
        for prec in ("fp32", "fp64"):
            r = time_rk4(prec, N, L, steps, 60, kernel_samples=0)
            b = (28 if prec == "fp32" else 56) * N * N
            out[f"{N}_{prec}"] = {"us_per_step": round(r["us_per_step"], 3),
                                  "hbm_frac": round(b / (r["us_per_step"] * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                                  "gcell_steps_per_s": round(N * N / (r["us_per_step"] * 1e-6) / 1e9, 2),
                                  "tile_rows": r["tile_rows"], "prefetch_rows": r["prefetch_rows"], "finite": r["finite"]}
    return out


def ensemble_inputs(i: int, N: int, amp: float, seed: int):
    """Instance i of the ensemble workload: dx = 0.5 um, Gaussian pump of width 10 um at amp * P_th,
    psi0 = 0.01 complex Gaussian noise from the instance's own seed, n0 = 0."""
    L = 0.5 * (N - 1)
    x = -L / 2 + 0.5 * np.arange(N)
    X, Y = np.meshgrid(x, x)
    P = (amp * PTH * np.exp(-(X ** 2 + Y ** 2) / (2 * 10.0 ** 2))).astype(np.float32)
    rng = np.random.default_rng(seed)
    psi = (0.01 * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))).astype(np.complex64)
    return psi, np.zeros((N, N), np.float32), P


def time_batched(instances: int, steps: int, seed: int = 99, graph_steps: int = 100, check: int = 8):
    """Ensemble of `instances` FP32 grids with sides drawn from {128,192,256,384,512} and pump amplitudes
    from U[0.8, 3] P_th: `steps` RK4 steps as one fused launch per step (hipGraph replay) against the same
    instances stepped one after the other through the plain single-grid entry point."""
    import torch
    from phoenix_amd import _lib as Lb
    rng = np.random.default_rng(seed)
    sides = rng.choice([128, 192, 256, 384, 512], size=instances)
    amps = rng.uniform(0.8, 3.0, size=instances)
    ph, _ = Lb.physics_defaults(512, 512, 255.5, 255.5)
    dt = 0.04
    stream = torch.cuda.Stream()
    ens = Lb.Ensemble("fp32", ph, [(int(N), int(N)) for N in sides], [(0.5, 0.5)] * instances)
    ens.set_stream(stream.cuda_stream)
    dev = torch.device("cuda:0")
    seq = []
    for i, N in enumerate(sides):
        psi, n, P = ensemble_inputs(i, int(N), float(amps[i]), 1000 + i)
        ens.set(i, "psi_plus", psi)
        ens.set(i, "n_plus", n)
        ens.set(i, "pump_plus", P)
        t = [torch.from_numpy(a).to(dev) for a in (psi, n, P)]
        seq.append((int(N), t[0], t[1], t[2], torch.empty_like(t[0]), torch.empty_like(t[1])))
    cells = int(sum(int(N) ** 2 for N in sides))
    torch.cuda.synchronize()
    # warm-up: autotune of the tile height and graph capture, without advancing the timed state twice
    ens.step(dt, graph_steps, graph_steps=graph_steps)
    ens.step(dt, graph_steps, graph_steps=graph_steps)
    ens.sync()
    for N, psi, n, P, bpsi, bn in seq:
        Lb.rk4_scalar_device(Lb.PHX_FP32, ph, N, N, (True, True), psi.data_ptr(), n.data_ptr(), bpsi.data_ptr(),
                             bn.data_ptr(), P.data_ptr(), None, dt, 2 * graph_steps, stream.cuda_stream)
    torch.cuda.synchronize()
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    e0.record(stream)
    ens.step(dt, steps, graph_steps=graph_steps)
    e1.record(stream)
    for N, psi, n, P, bpsi, bn in seq:
        Lb.rk4_scalar_device(Lb.PHX_FP32, ph, N, N, (True, True), psi.data_ptr(), n.data_ptr(), bpsi.data_ptr(),
                             bn.data_ptr(), P.data_ptr(), None, dt, steps, stream.cuda_stream)
    e2.record(stream)
    torch.cuda.synchronize()
    ms_b, ms_s = e0.elapsed_time(e1), e1.elapsed_time(e2)
    same = all(np.array_equal(ens.get(i, "psi_plus"), seq[i][1].cpu().numpy()) and
               np.array_equal(ens.get(i, "n_plus"), seq[i][2].cpu().numpy()) for i in range(min(check, instances)))
    ens.close()
    del seq
    torch.cuda.empty_cache()  # the next leg allocates GBs: give the cached blocks of this one back
    achieved = 28 * cells / (1e-3 * ms_b / steps) / 1e9
    return {"instances": instances, "steps": steps, "cells": cells, "sides": "{128,192,256,384,512} seed %d" % seed,
            "ms_batched": round(ms_b, 3), "ms_sequential": round(ms_s, 3), "speedup": round(ms_s / ms_b, 3),
            "us_per_step_batched": round(1e3 * ms_b / steps, 3),
            "gcell_steps_per_s": round(cells * steps / (1e-3 * ms_b) / 1e9, 2),
            "hbm_frac": round(achieved / HBM_PEAK_GBS, 4), "bitwise_equal_first_%d" % min(check, instances): same}


def smooth_random_field(rng, N: int, L: float, count: int, amp: tuple, width: tuple):
    """Sum of `count` Gaussians with centres uniform over the domain, widths uniform in `width`
    and amplitudes uniform in `amp`."""
    x = -L / 2 + (L / (N - 1)) * np.arange(N)
    X, Y = np.meshgrid(x, x)
    f = np.zeros((N, N))
    for _ in range(count):
        cx, cy = rng.uniform(-L / 2, L / 2, 2)
        w = rng.uniform(*width)
        f += rng.uniform(*amp) * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * w * w))
    return f


def time_adjoint(N: int = 512, steps: int = 2000, K: int = 50, batch: int = 16, seed: int = 77):
    """Gradient of J = sum_seeds int_target |psi(T)|^2 dV (target: disk of radius 20 um centred 60 um from the
    middle) with respect to P(x,y) and V(x,y), FP32, `batch` noise seeds (2000 + b): forward with a checkpoint
    every K steps, then the adjoint backward.  P and V are sums of 32 Gaussians (centres uniform over the
    domain, widths uniform in [5, 20] um, seed 77); P is scaled to a maximum of 1.5 P_th, V has amplitudes
    uniform in [-0.5, 0.5] meV.  Compared with the plain forward of the same batch."""
    import torch
    from phoenix_amd import _lib as Lb
    from phoenix_amd.adjoint import release_engines, rk4_evolve
    L = 0.5 * (N - 1)
    dt = 0.04
    rng = np.random.default_rng(seed)
    P = smooth_random_field(rng, N, L, 32, (0.0, 1.0), (5.0, 20.0))
    P *= 1.5 * PTH / P.max()
    V = smooth_random_field(rng, N, L, 32, (-5e-4, 5e-4), (5.0, 20.0))  # 1 code energy unit = 1 eV
    psi0 = np.stack([ensemble_inputs(0, N, 0.0, 2000 + b)[0] for b in range(batch)])
    x = -L / 2 + 0.5 * np.arange(N)
    X, Y = np.meshgrid(x, x)
    target = ((X - 60.0) ** 2 + Y ** 2 <= 20.0 ** 2).astype(np.float32)
    ph, _ = Lb.physics_defaults(N, N, L, L)
    dev = torch.device("cuda:0")
    tP = torch.tensor(P, dtype=torch.float32, device=dev, requires_grad=True)
    tV = torch.tensor(V, dtype=torch.float32, device=dev, requires_grad=True)
    tpsi = torch.tensor(psi0, dtype=torch.complex64, device=dev)
    tn = torch.zeros((batch, N, N), dtype=torch.float32, device=dev)
    tt = torch.tensor(target, dtype=torch.float32, device=dev)
    dV = 0.25

    def run():
        psiT, _ = rk4_evolve(tP, tV, tpsi, tn, physics=ph, dt=dt, steps=steps, checkpoint_every=K)
        return (psiT.abs() ** 2 * tt).sum() * dV

    J = run()  # warm-up: allocation of the checkpoint store
    J.backward()  # warm-up: tile autotune
    tP.grad = tV.grad = None
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ev[0].record()
    J = run()
    ev[1].record()
    J.backward()
    ev[2].record()
    torch.cuda.synchronize()
    ms_f, ms_b = ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2])
    gP, gV = tP.grad, tV.grad
    finite = bool(torch.isfinite(gP).all() and torch.isfinite(gV).all())
    # plain forward of the same batch: one uniform-batch RK4 launch per step, no checkpoints
    ens = Lb.Ensemble("fp32", ph, [(N, N)] * batch, [(0.5, 0.5)] * batch, reservoir=True, pump=True, potential=True)
    stream = torch.cuda.Stream()
    ens.set_stream(stream.cuda_stream)
    for b in range(batch):
        ens.set(b, "psi_plus", psi0[b])
        ens.set(b, "n_plus", np.zeros((N, N), np.float32))
        ens.set(b, "pump_plus", P.astype(np.float32))
        ens.set(b, "potential_plus", V.astype(np.float32))
    ens.step(dt, 100, graph_steps=100)
    torch.cuda.synchronize()
    ev[0].record(stream)
    ens.step(dt, steps, graph_steps=100)
    ev[1].record(stream)
    torch.cuda.synchronize()
    ms_plain = ev[0].elapsed_time(ev[1])
    ens.close()
    release_engines()
    return {"grid": f"{N}x{N}", "steps": steps, "checkpoint_every": K, "batch": batch, "objective": float(J),
            "ms_forward_checkpointed": round(ms_f, 2), "ms_backward": round(ms_b, 2), "ms_plain_forward": round(ms_plain, 2),
            "backward_over_forward": round(ms_b / ms_plain, 3),
            "gradient_over_forward": round((ms_f + ms_b) / ms_plain, 3), "finite": finite}


def leg_in_child(leg: str, *extra: str) -> dict:
    """One auxiliary leg of this script run in a fresh process.  The adjoint leg's host-launched
    forward and backward ran 40 % slower (82 vs 59 ms forward) in a process that had first run the
    multi-step RK4 legs and the 256-solver batched leg, and at full speed in a fresh process or after
    any part of that sequence; the child measures what a process running only the adjoint sees."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--leg", leg, *extra], capture_output=True, text=True,
                       timeout=900)
    if r.returncode != 0:
        return {"error": r.stderr[-500:]}
    return json.loads(r.stdout.strip().splitlines()[-1])[leg]


def cpu_baseline(N: int, L: float, steps: int = 200, warmup: int = 20):
    """The reference's CPU RK4 step restated in C (oracle/phx_cpu_rk4.c: padded subgrids with 4-cell
    halos, one halo sync per step, one OpenMP loop over subgrids running K1..K4 with `omp simd` rows;
    its FP64 build reproduces the FP64 oracle bit for bit), built -O3 -funroll-loops for the host's
    x86-64 level, on the host cores with OMP_PROC_BIND=true OMP_PLACES=cores.  Per precision the median
    of `steps` individually timed steps after `warmup`, best of the 4x4, 8x8, 16x16 and 32x32 subgrid
    splits of the 1024^2 workload.  Runs as a child process so that the OpenMP binding applies."""
    import subprocess
    env = dict(os.environ)
    cores = len(os.sched_getaffinity(0))
    env.setdefault("OMP_NUM_THREADS", str(cores))
    env["OMP_PROC_BIND"] = "true"
    env["OMP_PLACES"] = "cores"
    r = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "cpu_bench.py"), "--size", str(N), "--steps", str(steps),
                        "--warmup", str(warmup)], env=env, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        return {"error": r.stderr[-500:]}
    c = json.loads(r.stdout.strip().splitlines()[-1])
    return dict(value=c["fp32"]["best_us"], unit="µs/step", cores=c["threads"], kind="port",
                fp64_value=c["fp64"]["best_us"], per_split={"fp32": c["fp32"]["median_us_per_step"],
                                                            "fp64": c["fp64"]["median_us_per_step"]},
                fp64_256x256_1000_steps_ms=c["fp64_256_1000_steps_ms"],
                sample=f"the reference's CPU RK4 step restated in C (bitwise the FP64 oracle), {N}x{N} workload, "
                       f"FP32 (value) and FP64, median of {steps} steps after {warmup} warm-up, best of 4x4/8x8/16x16/32x32 "
                       f"subgrids (FP32 {c['fp32']['best_split']}, FP64 {c['fp64']['best_split']}); {c['threads']} OpenMP "
                       f"threads, OMP_PROC_BIND=true OMP_PLACES=cores, {c['cpu']}, {c['isa']}, -O3 -funroll-loops")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", choices=["fp32", "fp64"], default=None, help="time one precision only (profiling)")
    ap.add_argument("--kernel-samples", type=int, default=200, help="single-step launches timed with event pairs")
    ap.add_argument("--batch-instances", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--only-legs", action="store_true",
                    help="run only the batched-ensemble and adjoint legs (profiling); prints their objects")
    ap.add_argument("--leg", choices=["batched", "adjoint", "spinor", "fft", "large"], default=None,
                    help="run one leg only (profiling) and print its object")
    ap.add_argument("--large-sizes", default="4096,8192", help="grid sides of the large-grid leg (comma separated)")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-spinor", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="only run a plain device copy of this many MiB 10 times (calibrates rocprof byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); the path shards by halo exchange, "
              "which is designed but not run here.", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "no HIP device visible"
    if a.copy_mb > 0:
        src = torch.ones(a.copy_mb * 2 ** 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": src.numel() * 4, "launches": 10}))
        return
    if a.only_legs:
        print(json.dumps({"batched": time_batched(a.batch_instances, a.batch_steps),
                          "adjoint": time_adjoint(steps=a.adjoint_steps)}), flush=True)
        return
    large_sizes = [int(x) for x in a.large_sizes.split(",") if x]
    if a.leg is not None:
        legs = {"batched": lambda: time_batched(a.batch_instances, a.batch_steps),
                "adjoint": lambda: time_adjoint(steps=a.adjoint_steps),
                "spinor": lambda: time_spinor(steps=a.steps),
                "fft": lambda: {p: time_fft_filter(p, a.size, 0.5 * (a.size - 1)) for p in ("fp32", "fp64")},
                "large": lambda: time_large(large_sizes, a.steps)}
        print(json.dumps({a.leg: legs[a.leg]()}), flush=True)
        return
    N = a.size
    L = 0.5 * (N - 1)
    cells = N * N
    res = {}
    for prec in (["fp32", "fp64"] if a.only is None else [a.only]):
        res[prec] = time_rk4(prec, N, L, a.steps, a.warmup, a.kernel_samples)
    main_prec = "fp32" if "fp32" in res else "fp64"
    r = res[main_prec]
    bytes_per_cell = {"fp32": 28, "fp64": 56}
    # The timed region launches one multi-step kernel per up to 500 steps (or, with PHX_RK4_FLOW=0,
    # one kernel per step), back to back on one stream, so the
    # average launch duration over the region is total / steps (rocprof's per-kernel average agrees).
    kernel_s = r["us_per_step"] * 1e-6
    flow = r["multi_step_steps"] == a.steps
    per_launch = a.steps / max(r["multi_step_launches"], 1) if flow else 1
    traffic = measured_traffic(main_prec, r["tile_rows"], r["prefetch_rows"], "flow" if flow else "fused")
    if traffic is not None:
        traffic = round(traffic * per_launch)
    achieved = bytes_per_cell[main_prec] * cells / kernel_s / 1e9
    out = {
        "metric": METRIC,
        "value": round(r["us_per_step"], 4),
        "unit": "µs/step",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["us_per_step"] / 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": main_prec,
        "data": "synthetic",
        "config": {"workload": f"scalar GP+reservoir RK4, {N}x{N}, dx=0.5um, periodic, ring pump 1.5 P_th, dt=0.04ps",
                   "cells": cells, "parallelism": "single GPU"},
        "gcell_steps_per_s": round(cells / (r["us_per_step"] * 1e-6) / 1e9, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": f"{'rk4_scalar_flow' if flow else 'rk4_scalar_fused'}<{'float' if main_prec == 'fp32' else 'double'},"
                               f"TY={r['tile_rows']},PF={r['prefetch_rows']}>",
                     "steps_per_launch": per_launch,
                     "kernel_us": round(r["us_per_step"] * per_launch, 3),
                     "algorithmic_bytes": round(bytes_per_cell[main_prec] * cells * per_launch),
                     "bytes_per_cell": bytes_per_cell[main_prec]},
        "precisions": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                       for k, v in res.items()},
    }
    for k, v in res.items():
        out["precisions"][k]["roofline_frac"] = round(
            bytes_per_cell[k] * cells / (v["us_per_step"] * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)
    for k, v in res.items():
        out["precisions"][k]["fft_filter"] = time_fft_filter(k, N, L)
    if not a.no_batched and a.only is None:
        out["batched"] = time_batched(a.batch_instances, a.batch_steps)
    if not a.no_adjoint and a.only is None:
        out["adjoint"] = leg_in_child("adjoint", "--adjoint-steps", str(a.adjoint_steps))
    if not a.no_spinor and a.only is None:
        out["spinor_dp45"] = time_spinor()
    if not a.no_large and a.only is None:
        out["large_grids"] = time_large(large_sizes, 100)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(N, L)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
