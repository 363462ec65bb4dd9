"""GPU test: every launch shape the RK4 autotuners may choose gives the same result.

The one-step kernel (rk4_scalar_fused) is compiled for tile heights TY in {8, 10, 12, 16, 20, 24} and
prefetch depths PF in {2, 3, 4, 6, 8}, the multi-step kernel (rk4_scalar_flow) for TY in {8, ..., 20}
and the same PF; the autotuners pick among them by timing, so each must compute the same thing.
PHX_RK4_TY / PHX_RK4_PF force a shape.  Bar: bitwise equal to the default shape (same arithmetic in
every shape: explicit FMAs, no contraction), FP32 and FP64, a ragged zero-BC grid with a potential and
a periodic grid.  The default shape itself is compared with the oracle in test_gpu_rk4.py.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
PTH = 0.15 * 0.225 / 0.01
TYS = (8, 10, 12, 16, 20, 24)
FLOW_TYS = (8, 10, 12, 16, 20)
PFS = (2, 3, 4, 6, 8)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run(prec, N_c, N_r, bc, pot, steps):
    from phoenix_amd._lib import Solver
    L_x, L_y = 0.5 * (N_c - 1), 0.5 * (N_r - 1)
    args = ["--N", N_c, N_r, "--L", L_x, L_y, "--boundary", bc, bc, "--tstep", 0.04, "--tmax", 1e9,
            "--fftEvery", 1e12, "--outEvery", 1e12, "--pump", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide"]
    if pot:
        args += ["--potential", 0.02, "add", 8, 8, 3, -2, "both", 1, "none", "gauss+noDivide"]
    s = Solver([str(a) for a in args], prec)
    rng = np.random.default_rng(5)
    x = -L_x / 2 + 0.5 * np.arange(N_c)
    y = -L_y / 2 + 0.5 * np.arange(N_r)
    X, Y = np.meshgrid(x, y)
    s.set("psi_plus", 0.01 * (rng.standard_normal((N_r, N_c)) + 1j * rng.standard_normal((N_r, N_c))))
    s.set("n_plus", 3.0 * rng.random((N_r, N_c)))
    s.set("pump_plus", 1.5 * PTH * np.exp(-(X ** 2 + Y ** 2) / 400.0))
    assert s.step(steps) == steps
    out = (s.get("psi_plus"), s.get("n_plus"), s.info)
    s.close()
    return out


GRIDS = [("zero BC, potential, ragged", 200, 136, "zero", True), ("periodic", 256, 256, "periodic", False)]


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("name,N_c,N_r,bc,pot", GRIDS, ids=["zero_pot_200x136", "periodic_256"])
def test_every_one_step_shape_matches_default(monkeypatch, prec, name, N_c, N_r, bc, pot):
    steps = 30
    monkeypatch.setenv("PHX_RK4_FLOW", "0")
    ref_psi, ref_n, info = run(prec, N_c, N_r, bc, pot, steps)
    say(f"[tile shapes, one-step] {prec} {name}: default TY={info['tile_rows']} PF={info['prefetch_rows']}")
    bad = []
    for ty in TYS:
        for pf in PFS:
            monkeypatch.setenv("PHX_RK4_TY", str(ty))
            monkeypatch.setenv("PHX_RK4_PF", str(pf))
            psi, n, i = run(prec, N_c, N_r, bc, pot, steps)
            assert i["tile_rows"] == ty and i["prefetch_rows"] == pf
            if not (np.array_equal(psi, ref_psi) and np.array_equal(n, ref_n)):
                bad.append((ty, pf))
    say(f"  {len(TYS) * len(PFS)} shapes, differing: {bad}")
    assert not bad


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_every_multi_step_shape_matches_default(monkeypatch, prec):
    N_c, N_r, steps = 300, 200, 61
    monkeypatch.setenv("PHX_RK4_FLOW", "0")
    ref_psi, ref_n, _ = run(prec, N_c, N_r, "zero", True, steps)
    monkeypatch.setenv("PHX_RK4_FLOW", "1")
    say(f"[tile shapes, multi-step] {prec} {N_c}x{N_r} zero BC, potential, {steps} steps")
    bad = []
    for ty in FLOW_TYS:
        for pf in PFS:
            monkeypatch.setenv("PHX_RK4_TY", str(ty))
            monkeypatch.setenv("PHX_RK4_PF", str(pf))
            psi, n, i = run(prec, N_c, N_r, "zero", True, steps)
            # the first step runs alone; the other 60 are one multi-step launch
            assert i["multi_step_launches"] == 1 and i["multi_step_steps"] == steps - 1, (ty, pf, i)
            if not (np.array_equal(psi, ref_psi) and np.array_equal(n, ref_n)):
                bad.append((ty, pf))
    say(f"  {len(FLOW_TYS) * len(PFS)} shapes, differing: {bad}")
    assert not bad
