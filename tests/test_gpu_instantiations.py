"""Every instantiation of the fused Runge-Kutta kernel against the oracle: each tableau (fixed
and adaptive forms), each feature set the kernel is compiled for (TE-TM, reservoir, noise;
pump and potential are runtime inputs), FP64 and FP32, on a small non-square grid.

The adaptive forms run with dt_max = dt and a tolerance every attempt meets, so the step size
stays fixed and the comparison covers the error and norm reduction as well (the attempt count
must match the oracle's).  Bars: FP64 relative L2 <= 1e-12, FP32 <= 1e-5.
"""
import sys
import zlib

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = {"fp64": 1e-12, "fp32": 1e-5}
PTH = 0.15 * 0.225 / 0.01
CASES = [(k, False) for k in O.ITERATORS] + [(k, True) for k in O.ITERATORS if O.iter_info(k)[1] > 0]
FLAGS = [(t, r, n) for t in (False, True) for r in (False, True) for n in (False, True)]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("tetm,res,noise", FLAGS, ids=lambda v: str(int(v)))
@pytest.mark.parametrize("key,adaptive", CASES, ids=lambda v: str(v))
def test_instantiation_matches_oracle(key, adaptive, tetm, res, noise, prec):
    from phoenix_amd._lib import Solver
    Nc, Nr, L, dt, steps, seed = 40, 32, 19.5, 0.02, 4, 99
    Ly = L * (Nr - 1) / (Nc - 1)
    periodic = (res, not noise)  # alternate the boundary kinds over the feature sets
    say(f"[{key} {'adaptive' if adaptive else 'fixed'}] tetm={tetm} res={res} noise={noise} prec={prec}")
    rng = np.random.default_rng(zlib.crc32(f"{key}{adaptive}{tetm}{res}{noise}".encode()))
    dx = L / (Nc - 1)
    X, Y = np.meshgrid(-L / 2 + dx * np.arange(Nc), -Ly / 2 + dx * np.arange(Nr))
    P = 1.5 * PTH * np.exp(-0.5 * (X ** 2 + Y ** 2) / 5.0 ** 2)
    V = 2e-3 * np.exp(-0.5 * ((X - 2) ** 2 + (Y + 1) ** 2) / 3.0 ** 2)
    cplx = lambda: 0.05 * (rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc)))
    init = {"psi_plus": cplx(), "potential_plus": V}
    if res:
        init.update(n_plus=0.1 * rng.random((Nr, Nc)), pump_plus=P)
    if tetm:
        init.update(psi_minus=cplx(), potential_minus=V[::-1].copy())
        if res:
            init.update(n_minus=0.1 * rng.random((Nr, Nc)), pump_minus=0.7 * P)
    args = ["--N", Nc, Nr, "--L", L, Ly, "--boundary", "periodic" if periodic[0] else "zero",
            "periodic" if periodic[1] else "zero", "--outEvery", 1e9, "--fftEvery", 1e9,
            "--potential", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide", "--iterator", key,
            "--tstep", dt, "--tmax", (steps - 0.5) * dt]
    if res:
        args += ["--pump", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide"]
    if tetm:
        args += ["-tetm"]
    if noise:
        args += ["--dw", 0.05, "--initRandom", 0.0, seed]
    if adaptive:
        args += ["-adaptive", "--tol", 10.0, "--rkvdt", 1e-5, dt]
    s = Solver([str(a) for a in args], prec)
    for k, v in init.items():
        s.set(k, v)
    done = s.step(steps + 5)
    out = {k: s.get(k).astype(np.complex128 if k.startswith("psi") else np.float64) for k in init
           if k.startswith(("psi", "n_"))}
    info = s.info
    s.close()
    assert done == steps and info["iterator"] == key and info["adaptive"] == int(adaptive)
    p = O.params(Nc, Nr, L, Ly, periodic=periodic, tetm=tetm, reservoir=res, pump=res, potential=True,
                 stochastic=0.05 if noise else 0.0, noise_seed=seed)
    kw = dict(pump=init.get("pump_plus"), pot=V)
    if tetm:
        kw.update(psi_minus=init["psi_minus"], n_minus=init.get("n_minus"), pump_minus=init.get("pump_minus"),
                  pot_minus=init["potential_minus"])
    n0 = init.get("n_plus", np.zeros((Nr, Nc)))
    if tetm and not res:
        kw["n_minus"] = np.zeros((Nr, Nc))
    if adaptive:
        r = O.iter_adaptive_run(p, key, init["psi_plus"], n0, t=0.0, dt=dt, t_max=(steps - 0.5) * dt, tol=10.0,
                                dt_min=1e-5, dt_max=dt, **kw)
        assert r["accepted"] == steps and r["rejected"] == 0 and info["rejected"] == 0
        ref = {"psi_plus": r["psi"], "n_plus": r["n"], "psi_minus": r.get("psi_minus"), "n_minus": r.get("n_minus")}
    else:
        o = O.iter_run(p, key, init["psi_plus"], n0, dt=dt, steps=steps, **kw)
        ref = dict(zip(["psi_plus", "n_plus", "psi_minus", "n_minus"], o))
    errs = {k: rel_l2(v, ref[k]) for k, v in out.items()}
    say("  " + " ".join(f"{k} {e:.2e}" for k, e in errs.items()))
    assert max(errs.values()) <= TOL[prec]
