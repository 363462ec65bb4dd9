"""GPU tests of the multi-step RK4 launch (rk4_scalar_flow): phx_step runs stretches of fixed RK4
steps with constant sources (up to 500) as one launch in which every wavefront tile waits for its
eight neighbours' step counters.

Bar: bitwise equal to the same steps with PHX_RK4_FLOW=0 (blocks of 50 replayed from a hipGraph of
single-step launches, which the other GPU tests compare with the oracle), FP32 and FP64, periodic
and zero boundaries, ragged last tiles, one and two strips, with and without a potential.  The
time-out path is forced from the host with a spin limit of 0 polls: the error must reach
phx_sync, and every wave must still exit.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
PTH = 0.15 * 0.225 / 0.01


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def solver(prec, N_c, N_r, bc, pot, seed):
    from phoenix_amd._lib import Solver
    L_x, L_y = 0.5 * (N_c - 1), 0.5 * (N_r - 1)
    args = ["--N", N_c, N_r, "--L", L_x, L_y, "--boundary", bc, bc, "--tstep", 0.04, "--tmax", 1e9,
            "--fftEvery", 1e12, "--outEvery", 1e12, "--pump", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide"]
    if pot:
        args += ["--potential", 0.02, "add", 8, 8, 3, -2, "both", 1, "none", "gauss+noDivide"]
    s = Solver([str(a) for a in args], prec)
    rng = np.random.default_rng(seed)
    x = -L_x / 2 + 0.5 * np.arange(N_c)
    y = -L_y / 2 + 0.5 * np.arange(N_r)
    X, Y = np.meshgrid(x, y)
    s.set("psi_plus", 0.01 * (rng.standard_normal((N_r, N_c)) + 1j * rng.standard_normal((N_r, N_c))))
    s.set("n_plus", np.zeros((N_r, N_c)))
    s.set("pump_plus", 1.5 * PTH * np.exp(-(X ** 2 + Y ** 2) / 400.0))
    return s


CASES = [("bench 1024 periodic", 1024, 1024, "periodic", False),
         ("zero BC with potential", 300, 200, "zero", True),
         ("ragged periodic", 174, 134, "periodic", False),
         ("ragged zero BC", 174, 134, "zero", False),
         ("two strips, potential", 112, 64, "periodic", True)]


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("name,N_c,N_r,bc,pot", CASES, ids=[c[0].replace(" ", "_") for c in CASES])
def test_multi_step_launch_equals_step_by_step(monkeypatch, prec, name, N_c, N_r, bc, pot):
    steps = 301
    say(f"[multi-step] {name} {prec} {N_c}x{N_r} {bc} potential={pot}: {steps} steps")
    out = {}
    for flow in ("0", "1"):
        monkeypatch.setenv("PHX_RK4_FLOW", flow)
        s = solver(prec, N_c, N_r, bc, pot, 17)
        assert s.step(steps) == steps
        out[flow] = (s.get("psi_plus"), s.get("n_plus"), s.info)
        s.close()
    (p0, n0, i0), (p1, n1, i1) = out["0"], out["1"]
    say(f"  tile rows {i1['tile_rows']}, prefetch {i1['prefetch_rows']}: {i1['multi_step_steps']} steps in multi-step launches")
    assert i0["multi_step_steps"] == 0
    # the first step tunes the tile shape; the other 300 run as one launch
    assert i1["multi_step_steps"] == steps - 1 and i1["multi_step_launches"] == 1
    assert np.isfinite(p1).all() and np.isfinite(n1).all()
    assert np.array_equal(p0, p1) and np.array_equal(n0, n1)


def test_multi_step_wait_time_out_is_reported(monkeypatch):
    from phoenix_amd._lib import PhxError
    monkeypatch.setenv("PHX_RK4_FLOW", "1")
    monkeypatch.setenv("PHX_FLOW_SPIN_LIMIT", "0")
    s = solver("fp32", 1024, 1024, "periodic", False, 3)
    say("[multi-step] spin limit 0: 251 steps, the first neighbour wait that is not yet satisfied ends its wave")
    s.step(251)
    with pytest.raises(PhxError, match="waited too long"):
        s.sync()
    # the handle falls back to one launch per step afterwards (its counters no longer match)
    monkeypatch.delenv("PHX_FLOW_SPIN_LIMIT")
    before = s.info["multi_step_steps"]
    s.step(100)
    s.sync()
    assert s.info["multi_step_steps"] == before
    s.close()
