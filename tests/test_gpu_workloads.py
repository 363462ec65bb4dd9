"""GPU parity at the workload sizes themselves (not scaled-down stand-ins), against the FP64 oracle.

Bars: FP64 relative L2 <= 1e-12 (1e-11 for adaptive two-component runs, whose dt passes through
pow()), FP32 relative L2 <= 1e-5, after the step counts stated per test.  The oracle runs its
OpenMP subgrid loop on the host cores (about 4 ms per 1024^2 RK4 step on 16 cores).

  - 1024^2 scalar GP + reservoir, 4 x 4 subgrids on the oracle side, ring pump 1.5 P_th
    (radius 40 um, width 5 um), super-Gaussian k-space mask (exponent 6, radius 0.8 of the local
    coordinate) applied every 50 steps (--fftEvery 49.5 dt), 200 RK4 steps, FP32 and FP64;
  - 256^2, Gaussian pump (width 10 um, 1.5 P_th), 1000 RK4 steps, FP64 and FP32;
  - 2048^2 TE-TM spinor, adaptive Dormand-Prince 5(4), tolerance 1e-6, elliptic pump, without
    noise (--dw 0) and with it (--dw 1, the same Philox stream on both sides): the same accepted and
    rejected step counts, the same final dt, psi+- within the bar;
  - the 256-instance ensemble (sides from {128,...,512}, seed 99, pumps U[0.8, 3] P_th), 100 steps in
    one launch per step: eight instances (the smallest and largest among them) against the oracle;
  - 512^2 adjoint workload (P, V sums of 32 Gaussians, target disk of radius 20 um 60 um off centre,
    16 seeds): FP64 gradients against central differences at 64 random pixels after 2000 steps, and
    against the oracle's FP64 adjoint for one seed after 100 steps; on the same inputs, three FP64
    ensembles of different tile heights agree bit for bit over 50 steps.
"""
import math
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

PTH = 0.15 * 0.225 / 0.01
TOL = {"fp64": 1e-12, "fp32": 1e-5}
MASK_ARGS = ["--fftMask", "1.0", "add", "0.565685", "0.565685", "0", "0", "both", "6", "none", "gauss+noDivide+local"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def grid(N, L):
    x = -L / 2 + (L / (N - 1)) * np.arange(N)
    return np.meshgrid(x, x)


def rk4_solver(prec, N, L, steps, dt, extra=(), fft_every=1e9):
    from phoenix_amd._lib import Solver
    args = ["--N", N, N, "--subgrids", 4, 4, "--L", L, L, "--boundary", "periodic", "periodic", "--tstep", dt,
            "--tmax", (steps - 0.5) * dt, "--outEvery", 1e9, "--fftEvery", fft_every,
            "--pump", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide"] + list(extra)
    return Solver([str(a) for a in args], prec)


# ----------------------------------------------------------------------------- 1024^2, ring pump, filter

_RING = {}


def ring_case():
    """Inputs and the oracle's result of the 1024^2 ring-pump workload (computed once per module)."""
    if not _RING:
        N, L, dt, steps = 1024, 511.5, 0.04, 200
        X, Y = grid(N, L)
        r = np.hypot(X, Y)
        P = 1.5 * PTH * np.exp(-((r - 40.0) ** 2) / (2 * 5.0 ** 2))
        psi0 = O.gaussian_field(N, N, 0.01, 4242)
        dx = L / (N - 1)
        mask = O.envelope(N, N, L, L, dx, dx, 1.0, 0.565685, 0.565685, 0, 0, 6.0,
                          O.ENV_GAUSS | O.ENV_NODIVIDE | O.ENV_LOCAL).real
        p = O.params(N, N, L, L, pump=True, subgrids=(4, 4))
        say(f"[ring 1024] oracle on {O.set_threads(0)} threads, 4x4 subgrids, {steps} steps")
        o, on = psi0, np.zeros((N, N))
        for _ in range(steps // 50):
            o, on = O.rk4_run(p, o, on, pump=P, dt=dt, steps=50)
            o = O.fft_filter(o, mask)
        _RING.update(N=N, L=L, dt=dt, steps=steps, P=P, psi0=psi0, mask=mask, psi=o, n=on)
    return _RING


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_ring_1024_filter_every_50_steps_matches_oracle(prec):
    c = ring_case()
    say(f"[ring 1024] prec={prec}: {c['steps']} RK4 steps, k-space filter every 50 steps")
    s = rk4_solver(prec, c["N"], c["L"], c["steps"], c["dt"], MASK_ARGS, fft_every=49.5 * c["dt"])
    s.set("psi_plus", c["psi0"])
    s.set("n_plus", np.zeros_like(c["P"]))
    s.set("pump_plus", c["P"])
    m = s.get("fft_mask_plus").astype(np.float64)
    assert np.abs(m - c["mask"]).max() <= (1e-14 if prec == "fp64" else 1e-6)
    assert s.step(c["steps"] + 10) == c["steps"]
    g, gn = s.get("psi_plus").astype(np.complex128), s.get("n_plus").astype(np.float64)
    s.close()
    e, en = rel_l2(g, c["psi"]), rel_l2(gn, c["n"])
    say(f"  rel L2 psi {e:.3e} n {en:.3e}")
    assert e <= TOL[prec] and en <= TOL[prec]
    # the filter acts: the same 200 steps without it end elsewhere
    assert rel_l2(c["psi"], c["psi0"]) > 1e-2


# ----------------------------------------------------------------------------- 256^2, 1000 steps

@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_gauss_256_1000_steps_matches_oracle(prec):
    """FP64 after 1000 steps within 1e-12.  FP32 within 1e-5 after 200 steps (the bar of the 200-step
    comparison) and within 1e-4 after 1000: FP32 rounding accumulates (2.6e-5 measured at 1000 steps)."""
    N, L, dt, steps = 256, 127.5, 0.04, 1000
    say(f"[gauss 256] prec={prec}: {steps} RK4 steps")
    X, Y = grid(N, L)
    P = 1.5 * PTH * np.exp(-(X ** 2 + Y ** 2) / 200.0)
    psi0 = O.gaussian_field(N, N, 0.01, 1234)
    s = rk4_solver(prec, N, L, steps, dt)
    s.set("psi_plus", psi0)
    s.set("n_plus", np.zeros((N, N)))
    s.set("pump_plus", P)
    p = O.params(N, N, L, L, pump=True)
    o, on = psi0, np.zeros((N, N))
    for chunk, bar in ((200, TOL[prec]), (800, TOL[prec] if prec == "fp64" else 1e-4)):
        assert s.step(chunk) == chunk
        o, on = O.rk4_run(p, o, on, pump=P, dt=dt, steps=chunk)
        g, gn = s.get("psi_plus").astype(np.complex128), s.get("n_plus").astype(np.float64)
        e, en = rel_l2(g, o), rel_l2(gn, on)
        say(f"  after {s.info['iteration']} steps: rel L2 psi {e:.3e} n {en:.3e}; |psi|max {np.abs(o).max():.3e}")
        assert e <= bar and en <= bar
    assert s.info["iteration"] == steps
    s.close()


# ----------------------------------------------------------------------------- 2048^2 spinor DP45

@pytest.mark.parametrize("prec,dw,tmax", [("fp32", 0.0, 0.3), ("fp64", 0.0, 0.3), ("fp32", 1.0, 0.15)])
def test_spinor_2048_adaptive_matches_oracle(prec, dw, tmax):
    from phoenix_amd._lib import Solver
    N, L, dt0, seed = 2048, 1023.5, 0.04, 7
    say(f"[spinor 2048] prec={prec} dw={dw} tmax={tmax}")
    X, Y = grid(N, L)
    g = np.exp(-(X ** 2 + Y ** 2) / 200.0)
    ap = 1.5 * PTH * (1 + math.sin(math.pi / 4)) / 2
    am = 1.5 * PTH * (1 - math.sin(math.pi / 4)) / 2
    psi_p, psi_m = O.gaussian_field(N, N, 0.01, seed), O.gaussian_field(N, N, 0.01, seed + 1)
    zero = np.zeros((N, N))
    args = ["-tetm", "--N", N, N, "--L", L, L, "--boundary", "periodic", "periodic", "--iterator", "DP45",
            "-adaptive", "--tol", 1e-6, "--rkvdt", 1e-4, 0.1, "--tstep", dt0, "--tmax", tmax, "--outEvery", 1e9,
            "--fftEvery", 1e9, "--dw", dw, "--initRandom", 0.0, seed,
            "--pump", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide"]
    s = Solver([str(a) for a in args], prec)
    for k, v in (("psi_plus", psi_p), ("psi_minus", psi_m), ("n_plus", zero), ("n_minus", zero),
                 ("pump_plus", ap * g), ("pump_minus", am * g)):
        s.set(k, v)
    done = s.step(100000)
    info = s.info
    gp, gm = s.get("psi_plus").astype(np.complex128), s.get("psi_minus").astype(np.complex128)
    s.close()
    p = O.params(N, N, L, L, tetm=True, pump=True, stochastic=dw, noise_seed=seed)
    r = O.dp45_run(p, psi_p, zero, pump=ap * g, t=0.0, dt=dt0, t_max=tmax, tol=1e-6, dt_min=1e-4, dt_max=0.1,
                   psi_minus=psi_m, n_minus=zero, pump_minus=am * g)
    say(f"  gpu: {done} accepted {info['rejected']} rejected dt={info['dt']:.9e}; "
        f"oracle: {r['accepted']} accepted {r['rejected']} rejected dt={r['dt']:.9e}")
    assert done == r["accepted"] and info["rejected"] == r["rejected"]
    assert done + info["rejected"] >= 3
    e = max(rel_l2(gp, r["psi"]), rel_l2(gm, r["psi_minus"]))
    say(f"  rel L2 psi (max of components) {e:.3e}")
    if prec == "fp64":
        assert e <= 1e-11 and abs(info["dt"] - r["dt"]) <= 1e-12 * r["dt"]
    else:
        assert e <= 1e-5 and abs(info["dt"] - r["dt"]) <= 1e-5 * r["dt"]


# ----------------------------------------------------------------------------- 256-instance ensemble

@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_ensemble_256_instances_match_oracle(prec):
    from phoenix_amd import _lib as Lb
    instances, steps, dt = 256, 100, 0.04
    rng = np.random.default_rng(99)
    sides = rng.choice([128, 192, 256, 384, 512], size=instances)
    amps = rng.uniform(0.8, 3.0, size=instances)
    say(f"[ensemble 256] prec={prec}: {int(sum(int(n) ** 2 for n in sides))} cells, {steps} steps")
    ph, _ = Lb.physics_defaults(512, 512, 255.5, 255.5)
    ens = Lb.Ensemble(prec, ph, [(int(n), int(n)) for n in sides], [(0.5, 0.5)] * instances)
    inputs = []
    for i, N in enumerate(sides):
        N = int(N)
        X, Y = grid(N, 0.5 * (N - 1))
        P = amps[i] * PTH * np.exp(-(X ** 2 + Y ** 2) / (2 * 10.0 ** 2))
        psi = O.gaussian_field(N, N, 0.01, 1000 + i)
        ens.set(i, "psi_plus", psi)
        ens.set(i, "n_plus", np.zeros((N, N)))
        ens.set(i, "pump_plus", P)
        inputs.append((N, psi, P))
    ens.step(dt, steps, graph_steps=50)
    ens.sync()
    pick = sorted({int(np.argmin(sides)), int(np.argmax(sides)), 0, 1, 77, 128, 200, 255})
    worst = 0.0
    for i in pick:
        N, psi, P = inputs[i]
        g = ens.get(i, "psi_plus").astype(np.complex128)
        gn = ens.get(i, "n_plus").astype(np.float64)
        p = O.params(N, N, 0.5 * (N - 1), 0.5 * (N - 1), pump=True)
        o, on = O.rk4_run(p, psi, np.zeros((N, N)), pump=P, dt=dt, steps=steps)
        e = max(rel_l2(g, o), rel_l2(gn, on))
        say(f"  instance {i} ({N}^2, {amps[i]:.2f} P_th): rel L2 {e:.3e}")
        worst = max(worst, e)
    ens.close()
    assert worst <= TOL[prec]


# ----------------------------------------------------------------------------- 512^2 adjoint

def adjoint_workload(N=512, seed=77, batch=16):
    L = 0.5 * (N - 1)
    X, Y = grid(N, L)
    rng = np.random.default_rng(seed)

    def field(count, amp, width):
        f = np.zeros((N, N))
        for _ in range(count):
            cx, cy = rng.uniform(-L / 2, L / 2, 2)
            w = rng.uniform(*width)
            f += rng.uniform(*amp) * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * w * w))
        return f

    P = field(32, (0.0, 1.0), (5.0, 20.0))
    P *= 1.5 * PTH / P.max()
    V = field(32, (-5e-4, 5e-4), (5.0, 20.0))
    psi0 = np.stack([O.gaussian_field(N, N, 0.01, 2000 + b) for b in range(batch)])
    target = ((X - 60.0) ** 2 + Y ** 2 <= 20.0 ** 2).astype(float)
    return L, P, V, psi0, target


def adjoint_grads(N, L, P, V, psi0, target, steps, dt, K):
    import torch
    from phoenix_amd import _lib
    from phoenix_amd.adjoint import release_engines, rk4_evolve
    ph, _ = _lib.physics_defaults(N, N, L, L)
    dev = torch.device("cuda:0")
    tP = torch.tensor(P, dtype=torch.float64, device=dev, requires_grad=True)
    tV = torch.tensor(V, dtype=torch.float64, device=dev, requires_grad=True)
    tpsi = torch.tensor(psi0, dtype=torch.complex128, device=dev)
    tn = torch.zeros(psi0.shape, dtype=torch.float64, device=dev)
    tt = torch.tensor(target, dtype=torch.float64, device=dev)
    psiT, _ = rk4_evolve(tP, tV, tpsi, tn, physics=ph, dt=dt, steps=steps, checkpoint_every=K)
    J = (psiT.abs() ** 2 * tt).sum() * 0.25
    J.backward()
    torch.cuda.synchronize()
    out = float(J), tP.grad.cpu().numpy(), tV.grad.cpu().numpy()
    del tP, tV, tpsi, tn, psiT
    release_engines()
    return out


def test_fp64_ensembles_of_three_tile_heights_agree_bitwise():
    """Three FP64 ensembles (tile heights 8, 12 and 20, 16 grids of 512^2 each, reservoir, pump and
    potential) stepped side by side agree bit for bit after 1, 2, 5, 20 and 50 steps.  Regression:
    complex FP64 cells are written with 16-byte buffer stores; while their row offset sat in an SGPR
    soffset, nothing kept the next FMA from overwriting the store's data registers before the store
    had read them, and a few rows per step held -(gamma_r + R |psi|^2) as real part (the result of
    that FMA) in place of Re psi."""
    from phoenix_amd import _lib
    N, dt, B = 512, 0.04, 16
    L, P, V, psi0, _ = adjoint_workload(N, batch=B)
    ph, _ = _lib.physics_defaults(N, N, L, L)
    ens = []
    for ty in (8, 12, 20):
        e = _lib.Ensemble("fp64", ph, [(N, N)] * B, [(0.5, 0.5)] * B, reservoir=True, pump=True, potential=True)
        e.set_tile_rows(ty)
        for b in range(B):
            e.set(b, "psi_plus", psi0[b])
            e.set(b, "n_plus", np.zeros((N, N)))
            e.set(b, "pump_plus", P)
            e.set(b, "potential_plus", V)
        ens.append(e)
    done = 0
    for upto in (1, 2, 5, 20, 50):
        for e in ens:
            e.step(dt, upto - done)
        done = upto
        for e in ens:
            e.sync()
        for field in ("psi_plus", "n_plus"):
            ref = [ens[0].get(b, field) for b in range(B)]
            for k, ty in ((1, 12), (2, 20)):
                bad = [b for b in range(B) if not np.array_equal(ens[k].get(b, field), ref[b])]
                assert not bad, f"{field} after {upto} steps: tile height {ty} differs from 8 in grids {bad}"
        say(f"  {upto} steps: tile heights 8, 12, 20 agree bitwise")
    for e in ens:
        e.close()


def test_adjoint_512_gradients_match_finite_differences():
    from phoenix_amd import _lib
    N, dt, steps, K = 512, 0.04, 2000, 50
    L, P, V, psi0, target = adjoint_workload(N)
    B = psi0.shape[0]
    say(f"[adjoint 512] FP64, {B} seeds, {steps} steps, checkpoint every {K}")
    J0, gP, gV = adjoint_grads(N, L, P, V, psi0, target, steps, dt, K)
    ph, _ = _lib.physics_defaults(N, N, L, L)
    ens = _lib.Ensemble("fp64", ph, [(N, N)] * B, [(0.5, 0.5)] * B, reservoir=True, pump=True, potential=True)

    def J(Pv, Vv):
        for b in range(B):
            ens.set(b, "psi_plus", psi0[b])
            ens.set(b, "n_plus", np.zeros((N, N)))
            ens.set(b, "pump_plus", Pv)
            ens.set(b, "potential_plus", Vv)
        ens.step(dt, steps, graph_steps=100)
        ens.sync()
        return sum(float((np.abs(ens.get(b, "psi_plus")) ** 2 * target).sum()) for b in range(B)) * 0.25

    Jp = J(P, V)
    say(f"  J = {J0:.12e} (adjoint forward), {Jp:.12e} (ensemble forward)")
    assert abs(Jp - J0) <= 1e-12 * abs(J0)
    rng = np.random.default_rng(88)
    # pixels where the target's sensitivity is not negligible: inside a 60 um box around the disk
    X, Y = grid(N, L)
    near = np.argwhere((np.abs(X - 60.0) < 60.0) & (np.abs(Y) < 60.0))
    pix = near[rng.choice(len(near), 64, replace=False)]
    hP, hV = 1e-3 * np.abs(P).max(), 1e-3 * np.abs(V).max()
    worst, used = 0.0, 0
    for k, (r, c) in enumerate(pix):
        for g, base, h, is_P in ((gP, P, hP, True), (gV, V, hV, False)):
            a, b = base.copy(), base.copy()
            a[r, c] += h
            b[r, c] -= h
            fd = ((J(a, V) - J(b, V)) if is_P else (J(P, a) - J(P, b))) / (2 * h)
            if abs(fd) > 1e-3 * np.abs(g).max():
                worst = max(worst, abs(g[r, c] - fd) / abs(fd))
                used += 1
        if k % 16 == 15:
            say(f"  {k + 1} pixels, {used} well-conditioned differences, worst relative error {worst:.2e}")
    ens.close()
    say(f"  {used} well-conditioned differences, worst relative error {worst:.2e}")
    assert used >= 32
    assert worst <= 1e-4


def test_adjoint_512_matches_oracle_adjoint():
    N, dt, steps, K = 512, 0.04, 100, 50
    L, P, V, psi0, target = adjoint_workload(N, batch=1)
    say(f"[adjoint 512 vs oracle] FP64, 1 seed, {steps} steps")
    _, gP, gV = adjoint_grads(N, L, P, V, psi0, target, steps, dt, K)
    p = O.params(N, N, L, L, reservoir=True, pump=True, potential=True)
    oT, _ = O.rk4_run(p, psi0[0], np.zeros((N, N)), pump=P, pot=V, dt=dt, steps=steps)
    _, _, dP, dV, *_ = O.rk4_adjoint(p, psi0[0], np.zeros((N, N)), P, V, dt, steps, 2 * 0.25 * target * oT)
    e = max(rel_l2(gP, dP), rel_l2(gV, dV))
    say(f"  rel L2 dP {rel_l2(gP, dP):.2e} dV {rel_l2(gV, dV):.2e}")
    assert e <= 1e-10
