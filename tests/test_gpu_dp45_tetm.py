"""GPU parity of the generic fused RK kernel (phx_rk_generic.hip) against the FP64 oracle:
Dormand-Prince 5(4) fixed and adaptive, the TE-TM spinor model, and the stochastic term
(same Philox stream on both sides, so noisy runs are compared deterministically).

Bars: FP64 relative L2 <= 1e-12 (1e-11 for the 2-component adaptive runs, whose dt
sequence passes through pow()), FP32 relative L2 <= 1e-5.  Adaptive FP32 runs take
their dt from FP32 stage values, so their dt sequence differs from FP64 in the last
bits; they are compared with 1e-4 and must take the same number of steps.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
PTH = 0.15 * 0.225 / 0.01


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def inputs(N, L, seed, tetm=False):
    dx = L / (N - 1)
    x = -L / 2 + dx * np.arange(N)
    X, Y = np.meshgrid(x, x)
    P = 1.5 * PTH * np.exp(-0.5 * (X ** 2 + Y ** 2) / 6.0 ** 2)
    psi = O.gaussian_field(N, N, 0.01, seed)
    n = np.zeros((N, N))
    if not tetm:
        return psi, n, P
    chi = np.pi / 8
    Pp, Pm = P * (1 + np.sin(2 * chi)) / 2, P * (1 - np.sin(2 * chi)) / 2
    return psi, n, Pp, O.gaussian_field(N, N, 0.01, seed + 1), np.zeros((N, N)), Pm


def gpu(prec, N, L, args, fields, nsteps):
    from phoenix_amd._lib import Solver
    base = ["--N", N, N, "--L", L, L, "--boundary", "periodic", "periodic", "--outEvery", 1e9, "--fftEvery", 1e9,
            "--pump", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide"]
    s = Solver([str(a) for a in base + args], prec)
    for k, v in fields.items():
        s.set(k, v)
    done = s.step(nsteps)
    out = {k: s.get(k).astype(np.complex128 if k.startswith("psi") else np.float64) for k in fields if not k.startswith("pump")}
    info = s.info
    s.close()
    return out, done, info


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_dp45_fixed_step(prec):
    N, L, dt, steps = 64, 31.5, 0.04, 100
    say(f"[dp45 fixed] prec={prec}")
    psi, n, P = inputs(N, L, 11)
    out, done, _ = gpu(prec, N, L, ["--iterator", "DP45", "--tstep", dt, "--tmax", (steps - 0.5) * dt],
                       {"psi_plus": psi, "n_plus": n, "pump_plus": P}, steps + 5)
    assert done == steps
    p = O.params(N, N, L, L, pump=True)
    o, on = psi, n
    for k in range(steps):
        o, on, _, _ = O.dp45_attempt(p, o, on, pump=P, dt=dt, step=k)
    tol = 1e-12 if prec == "fp64" else 1e-5
    e, en = rel_l2(out["psi_plus"], o), rel_l2(out["n_plus"], on)
    say(f"  rel L2 psi {e:.3e} n {en:.3e}")
    assert e <= tol and en <= tol


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("tol", [1e-6, 1e-8])
def test_dp45_adaptive_trajectory(prec, tol):
    N, L, dt0, tmax = 64, 31.5, 0.04, 4.0
    say(f"[dp45 adaptive] prec={prec} tol={tol}")
    psi, n, P = inputs(N, L, 12)
    args = ["--iterator", "DP45", "-adaptive", "--tol", tol, "--rkvdt", 1e-4, 0.2, "--tstep", dt0, "--tmax", tmax]
    out, done, info = gpu(prec, N, L, args, {"psi_plus": psi, "n_plus": n, "pump_plus": P}, 100000)
    p = O.params(N, N, L, L, pump=True)
    r = O.dp45_run(p, psi, n, pump=P, t=0.0, dt=dt0, t_max=tmax, tol=tol, dt_min=1e-4, dt_max=0.2)
    say(f"  gpu: {done} steps, {info['rejected']} rejected, t={info['t']:.6f} dt={info['dt']:.6e}; "
        f"oracle: {r['accepted']} steps, {r['rejected']} rejected, t={r['t']:.6f} dt={r['dt']:.6e}")
    assert done == r["accepted"] and info["rejected"] == r["rejected"]
    e = rel_l2(out["psi_plus"], r["psi"])
    say(f"  rel L2 psi {e:.3e}")
    if prec == "fp64":
        assert e <= 1e-11 and abs(info["dt"] - r["dt"]) <= 1e-12 * r["dt"]
    else:
        assert e <= 1e-4


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("iterator", ["RK4", "DP45"])
def test_tetm_fixed_step(prec, iterator):
    N, L, dt, steps = 64, 31.5, 0.04, 100
    say(f"[tetm {iterator}] prec={prec}")
    psi, n, Pp, psim, nm, Pm = inputs(N, L, 21, tetm=True)
    fields = {"psi_plus": psi, "n_plus": n, "pump_plus": Pp, "psi_minus": psim, "n_minus": nm, "pump_minus": Pm}
    out, done, _ = gpu(prec, N, L, ["-tetm", "--iterator", iterator, "--tstep", dt, "--tmax", (steps - 0.5) * dt],
                       fields, steps + 5)
    assert done == steps
    p = O.params(N, N, L, L, tetm=True, pump=True)
    if iterator == "RK4":
        o, on, om, onm = O.rk4_run(p, psi, n, pump=Pp, dt=dt, steps=steps, psi_minus=psim, n_minus=nm, pump_minus=Pm)
    else:
        o, on, om, onm = psi, n, psim, nm
        for k in range(steps):
            o, on, _, _, om, onm = O.dp45_attempt(p, o, on, pump=Pp, dt=dt, step=k, psi_minus=om, n_minus=onm,
                                                  pump_minus=Pm)
    tol = 1e-12 if prec == "fp64" else 1e-5
    errs = [rel_l2(out["psi_plus"], o), rel_l2(out["psi_minus"], om), rel_l2(out["n_plus"], on),
            rel_l2(out["n_minus"], onm)]
    say(f"  rel L2 psi+ {errs[0]:.3e} psi- {errs[1]:.3e} n+ {errs[2]:.3e} n- {errs[3]:.3e}")
    assert max(errs) <= tol
    # the spin components must actually differ (TE-TM coupling and the elliptic pump are active)
    assert rel_l2(out["psi_plus"], out["psi_minus"]) > 1e-2


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_stochastic_rk4_deterministic_stream(prec):
    N, L, dt, steps, seed = 64, 31.5, 0.04, 50, 777
    say(f"[stochastic rk4] prec={prec}")
    psi, n, P = inputs(N, L, 13)
    args = ["--dw", 0.1, "--initRandom", 0.0, seed, "--tstep", dt, "--tmax", (steps - 0.5) * dt]
    out, done, _ = gpu(prec, N, L, args, {"psi_plus": psi, "n_plus": n, "pump_plus": P}, steps + 5)
    assert done == steps
    p = O.params(N, N, L, L, pump=True, stochastic=0.1, noise_seed=seed)
    o, on = O.rk4_run(p, psi, n, pump=P, dt=dt, steps=steps)
    tol = 1e-12 if prec == "fp64" else 1e-5
    e = rel_l2(out["psi_plus"], o)
    say(f"  rel L2 psi {e:.3e}")
    assert e <= tol
    # the noise is there: the same run without noise differs by far more than the tolerance
    p0 = O.params(N, N, L, L, pump=True)
    o0, _ = O.rk4_run(p0, psi, n, pump=P, dt=dt, steps=steps)
    assert rel_l2(o, o0) > 1e-2


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_spinor_adaptive_stochastic(prec):
    """The 2048^2 spinor workload's model on a small grid: -tetm, DP45 adaptive, noise, elliptic pump."""
    N, L, dt0, tmax, seed = 48, 23.5, 0.04, 1.0, 7
    say(f"[spinor adaptive stochastic] prec={prec}")
    psi, n, Pp, psim, nm, Pm = inputs(N, L, 31, tetm=True)
    fields = {"psi_plus": psi, "n_plus": n, "pump_plus": Pp, "psi_minus": psim, "n_minus": nm, "pump_minus": Pm}
    args = ["-tetm", "--iterator", "DP45", "-adaptive", "--tol", 1e-6, "--rkvdt", 1e-4, 0.1, "--tstep", dt0,
            "--tmax", tmax, "--dw", 0.05, "--initRandom", 0.0, seed]
    out, done, info = gpu(prec, N, L, args, fields, 100000)
    p = O.params(N, N, L, L, tetm=True, pump=True, stochastic=0.05, noise_seed=seed)
    r = O.dp45_run(p, psi, n, pump=Pp, t=0.0, dt=dt0, t_max=tmax, tol=1e-6, dt_min=1e-4, dt_max=0.1,
                   psi_minus=psim, n_minus=nm, pump_minus=Pm)
    say(f"  gpu: {done} steps {info['rejected']} rejected; oracle: {r['accepted']} steps {r['rejected']} rejected")
    assert done == r["accepted"]
    e = max(rel_l2(out["psi_plus"], r["psi"]), rel_l2(out["psi_minus"], r["psi_minus"]))
    say(f"  rel L2 psi (max of components) {e:.3e}")
    assert e <= (1e-11 if prec == "fp64" else 1e-4)


@pytest.mark.parametrize("pot", [False, True])
def test_spinor_tile_heights_agree(pot):
    """The FP32 TE-TM adaptive kernel runs 16- or 32-row tiles (PHX_GEN_TY, read at every launch):
    the same accepted and rejected counts, results equal to rounding (rel L2 <= 1e-6; the FP64 error sum
    is reduced over a different grouping of per-wave partials, so not bitwise).  The 32-row tiles come
    with and without the potential compiled in; both forms are checked against the 16-row kernel."""
    import os
    N, L, dt0, tmax, seed = 150, 74.5, 0.04, 0.6, 9
    psi, n, Pp, psim, nm, Pm = inputs(N, L, 41, tetm=True)
    fields = {"psi_plus": psi, "n_plus": n, "pump_plus": Pp, "psi_minus": psim, "n_minus": nm, "pump_minus": Pm}
    args = ["-tetm", "--iterator", "DP45", "-adaptive", "--tol", 1e-6, "--rkvdt", 1e-4, 0.1, "--tstep", dt0,
            "--tmax", tmax, "--dw", 0.05, "--initRandom", 0.0, seed]
    if pot:
        dx = L / (N - 1)
        X, Y = np.meshgrid(-L / 2 + dx * np.arange(N), -L / 2 + dx * np.arange(N))
        V = 2e-3 * np.exp(-0.5 * ((X - 3) ** 2 + (Y + 2) ** 2) / 6.0 ** 2)
        fields.update(potential_plus=V, potential_minus=V[::-1].copy())
        args += ["--potential", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide"]
    res = {}
    old = os.environ.get("PHX_GEN_TY")
    try:
        for ty in (16, 32):
            os.environ["PHX_GEN_TY"] = str(ty)
            say(f"[spinor tile height] TY={ty} potential={pot}")
            out, done, info = gpu("fp32", N, L, args, fields, 100000)
            res[ty] = (out, done, info["rejected"])
    finally:
        if old is None:
            os.environ.pop("PHX_GEN_TY", None)
        else:
            os.environ["PHX_GEN_TY"] = old
    for ty in (32,):
        assert res[ty][1] == res[16][1] and res[ty][2] == res[16][2]
        e = max(rel_l2(res[ty][0]["psi_plus"], res[16][0]["psi_plus"]), rel_l2(res[ty][0]["psi_minus"], res[16][0]["psi_minus"]))
        say(f"  TY={ty}: {res[ty][1]} steps {res[ty][2]} rejected, rel L2 vs TY=16 {e:.2e}")
        assert e <= 1e-6
