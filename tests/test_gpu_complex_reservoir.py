"""GPU parity of the complex reservoir (staged path, phx_staged.hip) against the oracle.

The reference stores n as Type::complex.  It becomes complex under a complex (iexp) temporal pump
amplitude (k_n += P a(w), kernel_gp_compute.cuh:70-75) and in imaginary time, where n advances by
-i dt like psi (cuda_macro.cuh:160-178) and is normalised after each step with psi
(solver_normalize_imag_time.cu:13-45).  The library then holds n as {re, im} cells and runs the step
on the staged path; phx_info.complex_reservoir reports it and the n fields are complex.

Bars: FP64 relative L2 <= 1e-12, FP32 <= 1e-5 (psi and n), against the oracle with
complex_reservoir set, the temporal amplitudes evaluated here with the envelope formula
(envelope.cpp:466-499) and handed to the oracle step by step.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
PTH = 0.15 * 0.225 / 0.01
TOL = {"fp64": 1e-12, "fp32": 1e-5}
GAUSS_NODIV = 1 | 8  # envelope type bits: gauss + noDivide
PUMP_T = (0.3, 0.5, 6.0)  # t0, sigma, freq of the pump's iexp envelope


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def t_iexp(t, t0, sigma, freq):
    return np.exp(-complex((t - t0) ** 2 / (2.0 * sigma * sigma), freq * (t - t0)))


def grid(Nc, Nr, L):
    dx = L / (Nc - 1)
    return dx, L * (Nr - 1) / (Nc - 1)


def pump_field(Nc, Nr, L, amp=1.5 * PTH):
    dx, Ly = grid(Nc, Nr, L)
    return amp * np.real(O.envelope(Nc, Nr, L, Ly, dx, dx, 1.0, 6.0, 6.0, 0.0, 0.0, 1.0, GAUSS_NODIV))


def start(Nc, Nr, seed, tetm):
    rng = np.random.default_rng(seed)
    out = {"psi_plus": 0.5 * (rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc))),
           "n_plus": rng.random((Nr, Nc)) + 0.2j * rng.standard_normal((Nr, Nc))}
    if tetm:
        out.update(psi_minus=0.5 * (rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc))),
                   n_minus=rng.random((Nr, Nc)) - 0.2j * rng.standard_normal((Nr, Nc)))
    return out


def run_gpu(prec, Nc, Nr, L, args, init, nsteps):
    from phoenix_amd._lib import Solver
    _, Ly = grid(Nc, Nr, L)
    base = ["--N", Nc, Nr, "--L", L, Ly, "--boundary", "periodic", "zero", "--outEvery", 1e9, "--fftEvery", 1e9]
    s = Solver([str(a) for a in base + args], prec)
    info = s.info
    assert info["complex_reservoir"] == 1 and info["staged"] == 1
    for k, v in init.items():
        s.set(k, v)
    done = s.step(nsteps)
    out = {k: s.get(k).astype(np.complex128) for k in init if k.startswith(("psi", "n_"))}
    s.close()
    return out, done


def oracle_steps(make_p, key, init, dt, steps, tetm, **fields):
    """`steps` single oracle steps, the parameters of step k from make_p(t_k); noise step index k."""
    psi, n = init["psi_plus"], init["n_plus"]
    pm, nm = init.get("psi_minus"), init.get("n_minus")
    t = 0.0
    for k in range(steps):
        r = O.iter_run(make_p(t), key, psi, n, dt=dt, steps=1, psi_minus=pm, n_minus=nm, step0=k, **fields)
        if tetm:
            psi, n, pm, nm = r
        else:
            psi, n = r
        t = t + dt
    return psi, n, pm, nm


def check(prec, out, o, tetm):
    errs = [rel_l2(out["psi_plus"], o[0]), rel_l2(out["n_plus"], o[1])]
    if tetm:
        errs += [rel_l2(out["psi_minus"], o[2]), rel_l2(out["n_minus"], o[3])]
    say(f"  rel L2 (psi+, n+[, psi-, n-]) {', '.join(f'{e:.2e}' for e in errs)}")
    assert max(errs) <= TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("key", ["RK4", "DP45"])
def test_complex_pump_envelope_matches_oracle(prec, key):
    """A pump with a complex (iexp) temporal envelope: n turns complex; scalar, periodic x / zero y."""
    Nc, Nr, L, dt, steps = 64, 48, 31.5, 0.04, 25
    say(f"[complex reservoir, iexp pump] {key} {prec} {Nc}x{Nr}")
    init = start(Nc, Nr, 31, False)
    P = pump_field(Nc, Nr, L)
    init["pump_plus"] = P
    args = ["--iterator", key, "--tstep", dt, "--tmax", (steps - 0.5) * dt,
            "--pump", 1.0, "add", 6.0, 6.0, 0, 0, "both", 1, "none", "gauss+noDivide", "time", "iexp", *PUMP_T]
    out, done = run_gpu(prec, Nc, Nr, L, args, init, steps + 5)
    assert done == steps

    def make_p(t):
        p = O.params(Nc, Nr, L, grid(Nc, Nr, L)[1], periodic=(True, False), pump=True)
        p.complex_reservoir = 1
        O.set_chain(p, "pump", [(P, None, t_iexp(t, *PUMP_T), t_iexp(t + dt, *PUMP_T))])
        return p

    o = oracle_steps(make_p, key, init, dt, steps, False)
    check(prec, out, o, False)
    # the pump's phase reached the reservoir: Im n moved well beyond the tolerance
    assert np.abs(o[1].imag - init["n_plus"].imag).max() > 1e-3


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_imaginary_time_with_reservoir_matches_oracle(prec):
    """--imagTime with a pump and a reservoir: psi and n advance by -i dt and are normalised."""
    Nc, Nr, L, dt, steps, amp = 64, 48, 31.5, 0.04, 25, 2.0
    say(f"[complex reservoir, imaginary time] {prec} {Nc}x{Nr}")
    init = start(Nc, Nr, 41, False)
    P = pump_field(Nc, Nr, L)
    init["pump_plus"] = P
    args = ["--imagTime", amp, "--tstep", dt, "--tmax", (steps - 0.5) * dt,
            "--pump", 1.0, "add", 6.0, 6.0, 0, 0, "both", 1, "none", "gauss+noDivide"]
    out, done = run_gpu(prec, Nc, Nr, L, args, init, steps + 5)
    assert done == steps

    def make_p(t):
        p = O.params(Nc, Nr, L, grid(Nc, Nr, L)[1], periodic=(True, False), pump=True)
        p.complex_reservoir = 1
        p.imag_time_amplitude = amp
        O.set_chain(p, "pump", [(P, None, 1.0, 1.0)])
        return p

    o = oracle_steps(make_p, "RK4", init, dt, steps, False)
    check(prec, out, o, False)
    dV = grid(Nc, Nr, L)[0] ** 2
    for f in (out["psi_plus"], out["n_plus"]):
        assert abs(np.sum(np.abs(f) ** 2) * dV - amp) < (1e-9 if prec == "fp64" else 1e-3) * amp


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_tetm_noise_complex_reservoir_matches_oracle(prec):
    """TE-TM, DP45 (fixed), noise with a complex reservoir (complex square root in the noise
    amplitude, the psi+ reservoir for both components), iexp pump on both components."""
    Nc, Nr, L, dt, steps, seed = 48, 40, 23.5, 0.04, 15, 321
    say(f"[complex reservoir, TE-TM + noise] {prec} {Nc}x{Nr}")
    init = start(Nc, Nr, 51, True)
    P = pump_field(Nc, Nr, L)
    init["pump_plus"] = P
    init["pump_minus"] = 0.5 * P
    args = ["-tetm", "--iterator", "DP45", "--dw", 0.05, "--initRandom", 0.0, seed, "--tstep", dt,
            "--tmax", (steps - 0.5) * dt,
            "--pump", 1.0, "add", 6.0, 6.0, 0, 0, "both", 1, "none", "gauss+noDivide", "time", "iexp", *PUMP_T]
    out, done = run_gpu(prec, Nc, Nr, L, args, init, steps + 5)
    assert done == steps

    def make_p(t):
        p = O.params(Nc, Nr, L, grid(Nc, Nr, L)[1], periodic=(True, False), tetm=True, pump=True, stochastic=0.05,
                     noise_seed=seed)
        p.complex_reservoir = 1
        O.set_chain(p, "pump", [(P, 0.5 * P, t_iexp(t, *PUMP_T), t_iexp(t + dt, *PUMP_T))])
        return p

    o = oracle_steps(make_p, "DP45", init, dt, steps, True)
    check(prec, out, o, True)
