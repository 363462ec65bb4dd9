"""GPU parity of imaginary-time propagation (--imagTime) against the oracle: psi advances by
-i dt in every stage sum and is normalised to sum |psi|^2 dV = amplitude after every step
(FP64 sums on the device, no host round trip).

Bars: FP64 relative L2 <= 1e-12, FP32 <= 1e-5; the norm after the run equals the amplitude.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = {"fp64": 1e-12, "fp32": 1e-5}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def trap(Nc, Nr, L):
    Ly = L * (Nr - 1) / (Nc - 1)
    dx = L / (Nc - 1)
    X, Y = np.meshgrid(-L / 2 + dx * np.arange(Nc), -Ly / 2 + dx * np.arange(Nr))
    return 2e-4 * (X ** 2 + 1.5 * Y ** 2), Ly


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("key,adaptive,tetm", [("RK4", False, False), ("Heun2", False, False), ("DP45", False, False),
                                               ("BoSha", True, False), ("RK4", False, True), ("Nystroem", False, True)])
def test_imaginary_time_matches_oracle(key, adaptive, tetm, prec):
    from phoenix_amd._lib import Solver
    Nc, Nr, L, dt, steps, amp = 64, 48, 31.5, 0.02, 40, 3.0
    V, Ly = trap(Nc, Nr, L)
    say(f"[imagTime {key} {'adaptive' if adaptive else 'fixed'} tetm={tetm}] prec={prec}")
    rng = np.random.default_rng(7)
    init = {"psi_plus": rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc)), "potential_plus": V}
    if tetm:
        init.update(psi_minus=rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc)),
                    potential_minus=V[::-1].copy())
    args = ["--N", Nc, Nr, "--L", L, Ly, "--boundary", "periodic", "zero", "--outEvery", 1e9, "--fftEvery", 1e9,
            "--potential", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide", "--imagTime", amp,
            "--iterator", key, "--tstep", dt, "--tmax", (steps - 0.5) * dt] + (["-tetm"] if tetm else [])
    if adaptive:
        args += ["-adaptive", "--tol", 1e-7, "--rkvdt", 1e-4, 0.05]
    s = Solver([str(a) for a in args], prec)
    for k, v in init.items():
        s.set(k, v)
    done = s.step(100000)
    out = {k: s.get(k).astype(np.complex128) for k in init if k.startswith("psi")}
    info = s.info
    s.close()
    p = O.params(Nc, Nr, L, Ly, periodic=(True, False), reservoir=False, potential=True, tetm=tetm,
                 imag_time_amplitude=amp)
    kw = dict(pot=V)
    if tetm:
        kw.update(psi_minus=init["psi_minus"], n_minus=np.zeros((Nr, Nc)), pot_minus=init["potential_minus"])
    if adaptive:
        r = O.iter_adaptive_run(p, key, init["psi_plus"], np.zeros((Nr, Nc)), t=0.0, dt=dt, t_max=(steps - 0.5) * dt,
                                tol=1e-7, dt_min=1e-4, dt_max=0.05, **kw)
        assert done == r["accepted"] and info["rejected"] == r["rejected"]
        ref = {"psi_plus": r["psi"], "psi_minus": r.get("psi_minus")}
    else:
        assert done == steps
        o = O.iter_run(p, key, init["psi_plus"], dt=dt, steps=steps, **kw)
        ref = {"psi_plus": o[0]}
        if tetm:
            ref["psi_minus"] = o[2]
    errs = {k: rel_l2(out[k], ref[k]) for k in out}
    say("  " + " ".join(f"{k} {e:.2e}" for k, e in errs.items()) + f" (steps {done})")
    assert max(errs.values()) <= TOL[prec]
    dV = p.dV
    for k in out:
        assert abs(np.sum(np.abs(out[k]) ** 2) * dV - amp) < (1e-12 if prec == "fp64" else 1e-5) * amp
