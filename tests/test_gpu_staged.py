"""GPU parity of the staged Runge-Kutta path (phx_staged.hip) against the oracle, through the
solver's C ABI.  The staged path runs what the fused kernels do not: pulses, several
non-constant temporal groups of a source, complex temporal envelopes of the potential, and
the 12- and 13-stage iterators DP85 and NSRK78.

Bars: FP64 relative L2 <= 1e-12, FP32 <= 1e-5 for fixed steps; adaptive runs take the same
number of accepted and rejected steps as the oracle and agree to 1e-11 (FP64).
Temporal amplitudes are evaluated here with the envelope formulas (envelope.cpp:466-499) and
handed to the oracle step by step; spatial fields are set through the C ABI or evaluated with
the oracle's envelope function.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
PTH = 0.15 * 0.225 / 0.01
TOL = {"fp64": 1e-12, "fp32": 1e-5}
GAUSS_NODIV = 1 | 8  # envelope type bits: gauss + noDivide


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def t_cos(t, t0, sigma, freq):
    p = (t - t0) / sigma
    return np.exp(-0.5 * p * p) * (1.0 + np.cos(freq * (t - t0))) / 2.0


def t_iexp(t, t0, sigma, freq):
    return np.exp(-complex((t - t0) ** 2 / (2.0 * sigma * sigma), freq * (t - t0)))


def grid_xy(Nc, Nr, L):
    dx = L / (Nc - 1)
    Ly = L * (Nr - 1) / (Nc - 1)
    return dx, Ly


def start(Nc, Nr, seed, tetm=False):
    rng = np.random.default_rng(seed)
    out = {"psi_plus": 0.05 * (rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc))),
           "n_plus": 0.1 * rng.random((Nr, Nc))}
    if tetm:
        out.update(psi_minus=0.05 * (rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc))),
                   n_minus=0.1 * rng.random((Nr, Nc)))
    return out


def env(Nc, Nr, L, amp, w, x, y, charge=0.0):
    dx, Ly = grid_xy(Nc, Nr, L)
    return O.envelope(Nc, Nr, L, Ly, dx, dx, amp, w, w, x, y, 1.0, GAUSS_NODIV, charge=charge)


def run_gpu(prec, Nc, Nr, L, args, init, nsteps, boundary=("periodic", "zero")):
    from phoenix_amd._lib import Solver
    _, Ly = grid_xy(Nc, Nr, L)
    base = ["--N", Nc, Nr, "--L", L, Ly, "--boundary", *boundary, "--outEvery", 1e9, "--fftEvery", 1e9]
    s = Solver([str(a) for a in base + args], prec)
    for k, v in init.items():
        s.set(k, v)
    done = s.step(nsteps)
    out = {k: s.get(k).astype(np.complex128 if k.startswith("psi") else np.float64) for k in init
           if k.startswith(("psi", "n_"))}
    info = s.info
    s.close()
    return out, done, info


def oracle_steps(p_of_t, key, init, dt, steps, tetm=False, **fields):
    """`steps` single oracle steps with the parameters p_of_t(t) of each step."""
    psi, n = init["psi_plus"], init["n_plus"]
    pm, nm = init.get("psi_minus"), init.get("n_minus")
    t = 0.0
    for _ in range(steps):
        p = p_of_t(t)
        r = O.iter_run(p, key, psi, n, dt=dt, steps=1, psi_minus=pm, n_minus=nm, **fields)
        if tetm:
            psi, n, pm, nm = r
        else:
            psi, n = r
        t = t + dt
    return psi, n, pm, nm


PULSE_T = (0.3, 0.4, 7.0)  # t0, sigma, freq of the pulse's iexp envelope


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("tetm", [False, True])
def test_pulse_matches_oracle(prec, tetm):
    """RK4 with a Gaussian pulse of topological charge 1 (complex field) and a complex (iexp)
    temporal envelope, TE-TM or scalar, periodic x / zero y."""
    Nc, Nr, L, dt, steps = 64, 48, 31.5, 0.04, 25
    say(f"[staged pulse] prec={prec} tetm={tetm} {Nc}x{Nr}")
    init = start(Nc, Nr, 11, tetm)
    P = 1.5 * PTH * np.real(env(Nc, Nr, L, 1.0, 6.0, 0.0, 0.0))
    init["pump_plus"] = P
    if tetm:
        init["pump_minus"] = 0.5 * P
    F = env(Nc, Nr, L, 1e-5, 5.0, 2.0, -1.0, charge=1.0)  # F/hbar ~ 0.015 / ps
    args = ["--tstep", dt, "--tmax", (steps - 0.5) * dt,
            "--pump", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide",
            "--pulse", 1e-5, "add", 5.0, 5.0, 2.0, -1.0, "both", 1, 1, "gauss+noDivide", "time", "iexp", *PULSE_T]
    if tetm:
        args = ["-tetm"] + args
    out, done, info = run_gpu(prec, Nc, Nr, L, args, init, steps + 5)
    assert done == steps and info["staged"] == 1

    def p_of_t(t):
        p = O.params(Nc, Nr, L, grid_xy(Nc, Nr, L)[1], periodic=(True, False), pump=True, tetm=tetm)
        O.set_chain(p, "pulse", [(F, F if tetm else None, t_iexp(t, *PULSE_T), t_iexp(t + dt, *PULSE_T))])
        return p

    o, on, om, onm = oracle_steps(p_of_t, "RK4", init, dt, steps, tetm, pump=P, pump_minus=0.5 * P if tetm else None)
    e, en = rel_l2(out["psi_plus"], o), rel_l2(out["n_plus"], on)
    say(f"  rel L2 psi {e:.3e} n {en:.3e}")
    assert e <= TOL[prec] and en <= TOL[prec]
    if tetm:
        assert rel_l2(out["psi_minus"], om) <= TOL[prec] and rel_l2(out["n_minus"], onm) <= TOL[prec]
    # the pulse matters
    p0 = O.params(Nc, Nr, L, grid_xy(Nc, Nr, L)[1], periodic=(True, False), pump=True, tetm=tetm)
    r0 = O.iter_run(p0, "RK4", init["psi_plus"], init["n_plus"], pump=P, dt=dt, steps=steps,
                    psi_minus=init.get("psi_minus"), n_minus=init.get("n_minus"), pump_minus=0.5 * P if tetm else None)
    assert rel_l2(r0[0], o) > 1e-3


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("key", ["RK4", "DP45", "Nystroem"])
def test_pump_and_potential_chains_match_oracle(prec, key):
    """Two pump groups (constant + cos temporal envelope) and two potential groups (constant +
    complex iexp envelope): one field per group, each at its own amplitude per stage."""
    Nc, Nr, L, dt, steps = 64, 48, 31.5, 0.04, 25
    tc, ti = (0.2, 0.5, 9.0), (0.4, 0.6, 5.0)
    say(f"[staged chains] {key} prec={prec}")
    init = start(Nc, Nr, 21)
    P0 = 1.5 * PTH * np.real(env(Nc, Nr, L, 1.0, 6.0, 0.0, 0.0))
    V0 = 2e-3 * np.real(env(Nc, Nr, L, 1.0, 4.0, 3.0, -2.0))
    init["pump_plus"] = P0  # group 0 (the constant group, first on the command line)
    init["potential_plus"] = V0
    P1 = np.real(env(Nc, Nr, L, 3.0, 4.0, -4.0, 3.0))
    V1 = np.real(env(Nc, Nr, L, 1e-3, 5.0, 1.0, 1.0))
    args = ["--iterator", key, "--tstep", dt, "--tmax", (steps - 0.5) * dt,
            "--pump", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide",
            "--pump", 3.0, "add", 4.0, 4.0, -4.0, 3.0, "both", 1, "none", "gauss+noDivide", "time", "cos", *tc,
            "--potential", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide",
            "--potential", 1e-3, "add", 5.0, 5.0, 1.0, 1.0, "both", 1, "none", "gauss+noDivide", "time", "iexp", *ti]
    out, done, info = run_gpu(prec, Nc, Nr, L, args, init, steps + 5)
    assert done == steps and info["staged"] == 1

    def p_of_t(t):
        p = O.params(Nc, Nr, L, grid_xy(Nc, Nr, L)[1], periodic=(True, False), pump=True, potential=True)
        O.set_chain(p, "pump", [(P0, None, 1.0, 1.0), (P1, None, t_cos(t, *tc), t_cos(t + dt, *tc))])
        O.set_chain(p, "potential", [(V0, None, 1.0, 1.0), (V1, None, t_iexp(t, *ti), t_iexp(t + dt, *ti))])
        return p

    o, on, _, _ = oracle_steps(p_of_t, key, init, dt, steps)
    e, en = rel_l2(out["psi_plus"], o), rel_l2(out["n_plus"], on)
    say(f"  rel L2 psi {e:.3e} n {en:.3e}")
    assert e <= TOL[prec] and en <= TOL[prec]


STAGED_KEYS = ["NSRK78", "DP85"]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("key", STAGED_KEYS)
def test_long_tableau_fixed_step_matches_oracle(prec, key):
    Nc, Nr, L, dt, steps = 72, 48, 35.5, 0.04, 20
    say(f"[staged iterator {key} fixed] prec={prec}")
    init = start(Nc, Nr, 31)
    P = 1.5 * PTH * np.real(env(Nc, Nr, L, 1.0, 6.0, 0.0, 0.0))
    V = 2e-3 * np.real(env(Nc, Nr, L, 1.0, 4.0, 3.0, -2.0))
    init.update(pump_plus=P, potential_plus=V)
    args = ["--iterator", key, "--tstep", dt, "--tmax", (steps - 0.5) * dt,
            "--pump", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide",
            "--potential", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide"]
    out, done, info = run_gpu(prec, Nc, Nr, L, args, init, steps + 5)
    assert done == steps and info["staged"] == 1 and info["iterator"] == key
    p = O.params(Nc, Nr, L, grid_xy(Nc, Nr, L)[1], periodic=(True, False), pump=True, potential=True)
    o, on = O.iter_run(p, key, init["psi_plus"], init["n_plus"], pump=P, pot=V, dt=dt, steps=steps)
    e, en = rel_l2(out["psi_plus"], o), rel_l2(out["n_plus"], on)
    say(f"  rel L2 psi {e:.3e} n {en:.3e}")
    assert e <= TOL[prec] and en <= TOL[prec]


@pytest.mark.parametrize("key", STAGED_KEYS)
def test_long_tableau_adaptive_trajectory_matches_oracle(key):
    Nc, Nr, L, dt0, tmax, tol = 64, 64, 31.5, 0.04, 2.0, 1e-9
    say(f"[staged iterator {key} adaptive] fp64 tol={tol}")
    init = start(Nc, Nr, 41)
    P = 1.5 * PTH * np.real(env(Nc, Nr, L, 1.0, 6.0, 0.0, 0.0))
    init["pump_plus"] = P
    args = ["--iterator", key, "-adaptive", "--tol", tol, "--rkvdt", 1e-4, 0.2, "--tstep", dt0, "--tmax", tmax,
            "--pump", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide"]
    out, done, info = run_gpu("fp64", Nc, Nr, L, args, init, 100000)
    p = O.params(Nc, Nr, L, L, periodic=(True, False), pump=True)
    r = O.iter_adaptive_run(p, key, init["psi_plus"], init["n_plus"], pump=P, t=0.0, dt=dt0, t_max=tmax, tol=tol,
                            dt_min=1e-4, dt_max=0.2)
    say(f"  gpu: {done} steps, {info['rejected']} rejected, t={info['t']:.6f}; oracle: {r['accepted']} steps, "
        f"{r['rejected']} rejected, t={r['t']:.6f}")
    assert done == r["accepted"] and info["rejected"] == r["rejected"]
    assert abs(info["t"] - r["t"]) < 1e-12 and abs(info["dt"] - r["dt"]) < 1e-11 * r["dt"]
    assert rel_l2(out["psi_plus"], r["psi"]) <= 1e-11 and rel_l2(out["n_plus"], r["n"]) <= 1e-11


@pytest.mark.parametrize("tetm", [False, True])
def test_staged_path_equals_fused_kernels(tetm):
    """A pulse of amplitude 0 routes the step through the staged path: the result equals the
    fused kernels' (scalar RK4, or TE-TM DP45 with noise) to rounding."""
    Nc, Nr, L, dt, steps = 64, 48, 31.5, 0.04, 20
    say(f"[staged vs fused] tetm={tetm}")
    init = start(Nc, Nr, 51, tetm)
    P = 1.5 * PTH * np.real(env(Nc, Nr, L, 1.0, 6.0, 0.0, 0.0))
    init["pump_plus"] = P
    args = ["--tstep", dt, "--tmax", (steps - 0.5) * dt, "--pump", 1.0, "add", 10, 10, 0, 0, "both", 1, "none",
            "gauss+noDivide"]
    if tetm:
        init["pump_minus"] = 0.7 * P
        args = ["-tetm", "--iterator", "DP45", "--dw", 0.5, "--initRandom", 0.0, 77] + args
    pulse0 = ["--pulse", 0.0, "add", 5.0, 5.0, 0, 0, "both", 1, "none", "gauss+noDivide", "time", "iexp", *PULSE_T]
    a, da, ia = run_gpu("fp64", Nc, Nr, L, args, init, steps + 5)
    b, db, ib = run_gpu("fp64", Nc, Nr, L, args + pulse0, init, steps + 5)
    assert da == db == steps and ia["staged"] == 0 and ib["staged"] == 1
    for k in a:
        e = rel_l2(b[k], a[k])
        say(f"  {k}: rel L2 {e:.3e}")
        assert e <= 1e-12
