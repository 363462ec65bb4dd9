"""GPU parity of the split-step Fourier iterator (--iterator SSFM / -ssfm) against the oracle
(orc_ssfm_run): LDS FFT round trips with the separable linear propagator, the pointwise
nonlinear step with the Euler reservoir update, and the noise step.

Bars: FP64 relative L2 <= 1e-12, FP32 <= 1e-5 (FP32 transforms of 64..256 points round at
about 1e-7 per pass).
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = {"fp64": 1e-12, "fp32": 1e-5}
PTH = 0.15 * 0.225 / 0.01


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("Nc,Nr,tetm,noise", [(64, 64, False, False), (128, 64, False, True), (64, 128, True, False),
                                             (64, 64, True, True), (60, 48, False, False), (50, 64, True, True)])
def test_ssfm_matches_oracle(Nc, Nr, tetm, noise, prec):
    from phoenix_amd._lib import Solver
    L, dt, steps, seed = 31.5, 0.02, 25, 31
    Ly = L * (Nr - 1) / (Nc - 1)
    say(f"[ssfm] {Nc}x{Nr} tetm={tetm} noise={noise} prec={prec}")
    dx = L / (Nc - 1)
    X, Y = np.meshgrid(-L / 2 + dx * np.arange(Nc), -Ly / 2 + dx * np.arange(Nr))
    P = 1.5 * PTH * np.exp(-0.5 * (X ** 2 + Y ** 2) / 6.0 ** 2)
    V = 2e-3 * np.exp(-0.5 * ((X - 3) ** 2 + (Y + 2) ** 2) / 4.0 ** 2)
    rng = np.random.default_rng(Nc + Nr + tetm)
    init = {"psi_plus": 0.05 * (rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc))),
            "n_plus": 0.1 * rng.random((Nr, Nc)), "pump_plus": P, "potential_plus": V}
    if tetm:
        init.update(psi_minus=0.05 * (rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc))),
                    n_minus=0.1 * rng.random((Nr, Nc)), pump_minus=0.6 * P, potential_minus=V[::-1].copy())
    args = ["--N", Nc, Nr, "--L", L, Ly, "--boundary", "periodic", "periodic", "--outEvery", 1e9, "--fftEvery", 1e9,
            "--pump", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide",
            "--potential", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide",
            "-ssfm" if noise else "--iterator", *([] if noise else ["SSFM"]), "--tstep", dt, "--tmax", (steps - 0.5) * dt]
    if tetm:
        args += ["-tetm"]
    if noise:
        args += ["--dw", 0.05, "--initRandom", 0.0, seed]
    s = Solver([str(a) for a in args], prec)
    for k, v in init.items():
        s.set(k, v)
    done = s.step(steps + 5)
    out = {k: s.get(k).astype(np.complex128 if k.startswith("psi") else np.float64) for k in init
           if k.startswith(("psi", "n_"))}
    info = s.info
    s.close()
    assert done == steps and info["iterator"] == "SSFM"
    p = O.params(Nc, Nr, L, Ly, tetm=tetm, pump=True, potential=True, stochastic=0.05 if noise else 0.0,
                 noise_seed=seed)
    kw = dict(pump=P, pot=V)
    if tetm:
        kw.update(psi_minus=init["psi_minus"], n_minus=init["n_minus"], pump_minus=init["pump_minus"],
                  pot_minus=init["potential_minus"])
    o = O.ssfm_run(p, init["psi_plus"], init["n_plus"], dt=dt, steps=steps, **kw)
    ref = dict(zip(["psi_plus", "n_plus", "psi_minus", "n_minus"], o))
    errs = {k: rel_l2(out[k], ref[k]) for k in out}
    say("  " + " ".join(f"{k} {e:.2e}" for k, e in errs.items()))
    assert max(errs.values()) <= TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("tetm", [False, True])
def test_ssfm_imaginary_time_with_reservoir_matches_oracle(prec, tetm):
    """--imagTime with the split-step iterator: the step takes no -i dt, then psi and the (real)
    reservoir are rescaled to the amplitude after every step (solver_normalize_imag_time.cu:13-45)."""
    from phoenix_amd._lib import Solver
    Nc, Nr, L, dt, steps, amp = 64, 48, 31.5, 0.02, 25, 2.0
    Ly = L * (Nr - 1) / (Nc - 1)
    say(f"[ssfm imaginary time] tetm={tetm} prec={prec}")
    dx = L / (Nc - 1)
    X, Y = np.meshgrid(-L / 2 + dx * np.arange(Nc), -Ly / 2 + dx * np.arange(Nr))
    P = 1.5 * PTH * np.exp(-0.5 * (X ** 2 + Y ** 2) / 6.0 ** 2)
    rng = np.random.default_rng(5)
    init = {"psi_plus": 0.5 * (rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc))),
            "n_plus": rng.random((Nr, Nc)), "pump_plus": P}
    if tetm:
        init.update(psi_minus=0.5 * (rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc))),
                    n_minus=rng.random((Nr, Nc)), pump_minus=0.6 * P)
    args = ["--N", Nc, Nr, "--L", L, Ly, "--boundary", "periodic", "periodic", "--outEvery", 1e9, "--fftEvery", 1e9,
            "--pump", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide", "--iterator", "SSFM",
            "--imagTime", amp, "--tstep", dt, "--tmax", (steps - 0.5) * dt] + (["-tetm"] if tetm else [])
    s = Solver([str(a) for a in args], prec)
    assert s.info["complex_reservoir"] == 0
    for k, v in init.items():
        s.set(k, v)
    done = s.step(steps + 5)
    out = {k: s.get(k).astype(np.complex128 if k.startswith("psi") else np.float64) for k in init
           if k.startswith(("psi", "n_"))}
    s.close()
    assert done == steps
    p = O.params(Nc, Nr, L, Ly, tetm=tetm, pump=True)
    p.imag_time_amplitude = amp
    kw = dict(pump=P)
    if tetm:
        kw.update(psi_minus=init["psi_minus"], n_minus=init["n_minus"], pump_minus=init["pump_minus"])
    o = O.ssfm_run(p, init["psi_plus"], init["n_plus"], dt=dt, steps=steps, **kw)
    ref = dict(zip(["psi_plus", "n_plus", "psi_minus", "n_minus"], o))
    errs = {k: rel_l2(out[k], ref[k]) for k in out}
    say("  " + " ".join(f"{k} {e:.2e}" for k, e in errs.items()))
    assert max(errs.values()) <= TOL[prec]
    for k in out:
        assert abs(np.sum(np.abs(out[k]) ** 2) * dx * dx - amp) < (1e-9 if prec == "fp64" else 1e-3) * amp


def _env(Nc, Nr, L, amp, w, x, y, charge=0.0):
    dx = L / (Nc - 1)
    Ly = L * (Nr - 1) / (Nc - 1)
    return O.envelope(Nc, Nr, L, Ly, dx, dx, amp, w, w, x, y, 1.0, 1 | 8, charge=charge)  # gauss + noDivide


def _t_cos(t, t0, sigma, freq):
    q = (t - t0) / sigma
    return np.exp(-0.5 * q * q) * (1.0 + np.cos(freq * (t - t0))) / 2.0


def _t_iexp(t, t0, sigma, freq):
    return np.exp(-complex((t - t0) ** 2 / (2.0 * sigma * sigma), freq * (t - t0)))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("tetm", [False, True])
def test_ssfm_source_chains_and_pulse_match_oracle(prec, tetm):
    """SSFM with two pump groups (constant + cos envelope), two potential groups (constant + complex
    iexp envelope) and a pulse of charge 1 with an iexp envelope: each group at its amplitude of the
    step start; the pulse enters the independent step as (dt/hbar) F a (scalar) or -i (dt/hbar) F a
    (TE-TM)."""
    from phoenix_amd._lib import Solver
    Nc, Nr, L, dt, steps = 64, 64, 31.5, 0.02, 20
    tc, ti, tp = (0.1, 0.3, 9.0), (0.2, 0.4, 5.0), (0.15, 0.3, 7.0)
    say(f"[ssfm chains] tetm={tetm} prec={prec}")
    rng = np.random.default_rng(61 + tetm)
    P0 = 1.5 * PTH * np.real(_env(Nc, Nr, L, 1.0, 6.0, 0.0, 0.0))
    V0 = 2e-3 * np.real(_env(Nc, Nr, L, 1.0, 4.0, 3.0, -2.0))
    P1 = np.real(_env(Nc, Nr, L, 3.0, 4.0, -4.0, 3.0))
    V1 = np.real(_env(Nc, Nr, L, 1e-3, 5.0, 1.0, 1.0))
    F = _env(Nc, Nr, L, 1e-5, 5.0, 2.0, -1.0, charge=1.0)
    init = {"psi_plus": 0.05 * (rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc))),
            "n_plus": 0.1 * rng.random((Nr, Nc)), "pump_plus": P0, "potential_plus": V0}
    if tetm:
        init.update(psi_minus=0.05 * (rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc))),
                    n_minus=0.1 * rng.random((Nr, Nc)), pump_minus=0.5 * P0, potential_minus=V0[::-1].copy())
    args = ["--N", Nc, Nr, "--L", L, L, "--boundary", "periodic", "periodic", "--outEvery", 1e9, "--fftEvery", 1e9,
            "--iterator", "SSFM", "--tstep", dt, "--tmax", (steps - 0.5) * dt,
            "--pump", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide",
            "--pump", 3.0, "add", 4.0, 4.0, -4.0, 3.0, "both", 1, "none", "gauss+noDivide", "time", "cos", *tc,
            "--potential", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide",
            "--potential", 1e-3, "add", 5.0, 5.0, 1.0, 1.0, "both", 1, "none", "gauss+noDivide", "time", "iexp", *ti,
            "--pulse", 1e-5, "add", 5.0, 5.0, 2.0, -1.0, "both", 1, 1, "gauss+noDivide", "time", "iexp", *tp]
    if tetm:
        args += ["-tetm"]
    s = Solver([str(a) for a in args], prec)
    for k, v in init.items():
        s.set(k, v)
    done = s.step(steps + 5)
    out = {k: s.get(k).astype(np.complex128 if k.startswith("psi") else np.float64) for k in init
           if k.startswith(("psi", "n_"))}
    info = s.info
    s.close()
    assert done == steps and info["iterator"] == "SSFM" and info["staged"] == 0
    psi, n = init["psi_plus"], init["n_plus"]
    pm, nm = init.get("psi_minus"), init.get("n_minus")
    pump_m = 0.5 * P0 if tetm else None
    pot_m = init.get("potential_minus")
    t = 0.0
    for st in range(steps):
        p = O.params(Nc, Nr, L, L, tetm=tetm, pump=True, potential=True)
        O.set_chain(p, "pump", [(P0, pump_m, 1.0, 1.0), (P1, P1 if tetm else None, _t_cos(t, *tc), 0.0)])
        O.set_chain(p, "potential", [(V0, pot_m, 1.0, 1.0), (V1, V1 if tetm else None, _t_iexp(t, *ti), 0.0)])
        O.set_chain(p, "pulse", [(F, F if tetm else None, _t_iexp(t, *tp), 0.0)])
        r = O.ssfm_run(p, psi, n, dt=dt, steps=1, psi_minus=pm, n_minus=nm, step0=st)
        if tetm:
            psi, n, pm, nm = r
        else:
            psi, n = r
        t = t + dt
    ref = {"psi_plus": psi, "n_plus": n, "psi_minus": pm, "n_minus": nm}
    errs = {k: rel_l2(out[k], ref[k]) for k in out}
    say("  " + " ".join(f"{k} {e:.2e}" for k, e in errs.items()))
    assert max(errs.values()) <= TOL[prec]
    # the pulse and the complex potential amplitude matter at this size
    p = O.params(Nc, Nr, L, L, tetm=tetm, pump=True, potential=True)
    r0 = O.ssfm_run(p, init["psi_plus"], init["n_plus"], pump=P0, pot=V0, dt=dt, steps=steps,
                    psi_minus=init.get("psi_minus"), n_minus=init.get("n_minus"), pump_minus=pump_m, pot_minus=pot_m)
    assert rel_l2(r0[0], psi) > 1e-4


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("tetm,noise,imag", [(False, False, 0.0), (False, True, 0.0), (True, True, 0.0),
                                             (False, False, 2.0)])
def test_ssfm_complex_pump_envelope_matches_oracle(prec, tetm, noise, imag):
    """SSFM with a pump of complex (iexp) temporal envelope: the reservoir is complex ({re, im} cells),
    its Euler step takes P a whole, Im n enters the nonlinear exponent, the noise amplitude is the
    principal complex square root; with --imagTime psi and the complex reservoir are normalised after
    every step."""
    from phoenix_amd._lib import Solver
    Nc, Nr, L, dt, steps, seed = 64, 48, 31.5, 0.02, 25, 77
    tp = (0.2, 0.4, 6.0)
    Ly = L * (Nr - 1) / (Nc - 1)
    say(f"[ssfm complex reservoir] tetm={tetm} noise={noise} imag={imag} prec={prec}")
    P = 1.5 * PTH * np.real(_env(Nc, Nr, L, 1.0, 6.0, 0.0, 0.0))
    rng = np.random.default_rng(91 + 2 * tetm + noise)
    init = {"psi_plus": 0.3 * (rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc))),
            "n_plus": rng.random((Nr, Nc)) + 0.2j * rng.standard_normal((Nr, Nc)), "pump_plus": P}
    if tetm:
        init.update(psi_minus=0.3 * (rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc))),
                    n_minus=rng.random((Nr, Nc)) - 0.2j * rng.standard_normal((Nr, Nc)), pump_minus=0.5 * P)
    args = ["--N", Nc, Nr, "--L", L, Ly, "--boundary", "periodic", "periodic", "--outEvery", 1e9, "--fftEvery", 1e9,
            "--iterator", "SSFM", "--tstep", dt, "--tmax", (steps - 0.5) * dt,
            "--pump", 1.0, "add", 6.0, 6.0, 0, 0, "both", 1, "none", "gauss+noDivide", "time", "iexp", *tp]
    if tetm:
        args += ["-tetm"]
    if noise:
        args += ["--dw", 0.05, "--initRandom", 0.0, seed]
    if imag:
        args += ["--imagTime", imag]
    s = Solver([str(a) for a in args], prec)
    info = s.info
    assert info["complex_reservoir"] == 1 and info["iterator"] == "SSFM"
    for k, v in init.items():
        s.set(k, v)
    done = s.step(steps + 5)
    out = {k: s.get(k).astype(np.complex128) for k in init if k.startswith(("psi", "n_"))}
    assert s.info["staged"] == 0
    s.close()
    assert done == steps
    psi, n = init["psi_plus"], init["n_plus"]
    pm, nm = init.get("psi_minus"), init.get("n_minus")
    t = 0.0
    for st in range(steps):
        p = O.params(Nc, Nr, L, Ly, tetm=tetm, pump=True, stochastic=0.05 if noise else 0.0, noise_seed=seed)
        p.complex_reservoir = 1
        p.imag_time_amplitude = imag
        O.set_chain(p, "pump", [(P, 0.5 * P if tetm else None, _t_iexp(t, *tp), 0.0)])
        r = O.ssfm_run(p, psi, n, dt=dt, steps=1, psi_minus=pm, n_minus=nm, step0=st)
        if tetm:
            psi, n, pm, nm = r
        else:
            psi, n = r
        t = t + dt
    ref = {"psi_plus": psi, "n_plus": n, "psi_minus": pm, "n_minus": nm}
    errs = {k: rel_l2(out[k], ref[k]) for k in out}
    say("  " + " ".join(f"{k} {e:.2e}" for k, e in errs.items()))
    assert max(errs.values()) <= TOL[prec]
    if not tetm:  # the pump's phase reached the reservoir (TE-TM reservoirs do not evolve under SSFM)
        assert np.abs(n.imag - init["n_plus"].imag).max() > 1e-3
    if imag:
        dV = (L / (Nc - 1)) ** 2
        for k in out:
            assert abs(np.sum(np.abs(out[k]) ** 2) * dV - imag) < (1e-9 if prec == "fp64" else 1e-3) * imag
