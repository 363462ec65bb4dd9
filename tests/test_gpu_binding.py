"""GPU tests of the main-loop and Solver-subclass surfaces: phx_step_at (one Solver::step without
iterate()'s bookkeeping), the main loop's landing step (phx_run_interval, main.cu:84-101) and
--initRandom (solver_initialization.cu:71-79).

Bars: phx_step_at driven by a host restatement of iterate() equals phx_step bit for bit (RK4,
adaptive DP45 with noise); landing-step runs: FP64 rel L2 <= 1e-12, FP32 <= 1e-5 against the
oracle stepped with the same dt sequence, the same t at every output time; --initRandom: the
initial fields equal the oracle's restatement (pinned against libstdc++) bit for bit, and 100 RK4
steps from them stay within the parity bar.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
PTH = 0.15 * 0.225 / 0.01
TOL = {"fp64": 1e-12, "fp32": 1e-5}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def pump(N, L):
    x = -L / 2 + (L / (N - 1)) * np.arange(N)
    X, Y = np.meshgrid(x, x)
    return 1.5 * PTH * np.exp(-(X ** 2 + Y ** 2) / (2 * 6.0 ** 2))


def solver(prec, N, L, extra, tmax, dt=0.04, out_every=1e9):
    from phoenix_amd._lib import Solver
    args = ["--N", N, N, "--L", L, L, "--boundary", "periodic", "periodic", "--tstep", dt, "--tmax", tmax,
            "--outEvery", out_every, "--fftEvery", 1e9,
            "--pump", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide"] + list(extra)
    return Solver([str(a) for a in args], prec)


def rnd(prec):
    return (lambda v: float(np.float32(v))) if prec == "fp32" else float


# ----------------------------------------------------------------------------- phx_step_at

@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("kind", ["rk4", "dp45_adaptive_noise"])
def test_step_at_equals_iterate(prec, kind):
    N, L, steps = 64, 31.5, 30
    say(f"[step_at] prec={prec} kind={kind}")
    P = pump(N, L)
    psi0 = O.gaussian_field(N, N, 0.01, 5)
    extra = [] if kind == "rk4" else ["--iterator", "DP45", "-adaptive", "--tol", 1e-9, "--rkvdt", 1e-4, 0.2,
                                      "--dw", 0.05, "--initRandom", 0.0, 99]
    a = solver(prec, N, L, extra, tmax=1e9)
    b = solver(prec, N, L, extra, tmax=1e9)
    for s in (a, b):
        s.set("psi_plus", psi0)
        s.set("n_plus", np.zeros((N, N)))
        s.set("pump_plus", P)
    for _ in range(steps):
        assert a.step(1) == 1
    r = rnd(prec)
    t, dt, rej = 0.0, b.time()[1], 0
    for _ in range(steps):
        t_pre = t
        t = r(t + dt)  # iterate(): t += dt before the step (solver.cu:50)
        dt, k = b.step_at(t_pre, dt, variable_time_step=True)
        rej += k
    assert b.time()[0] == 0.0  # step_at leaves the handle's t alone
    ia = a.info
    say(f"  iterate: t={ia['t']:.9g} dt={ia['dt']:.9g} rejected={ia['rejected']}; "
        f"step_at: t={t:.9g} dt={dt:.9g} rejected={rej}")
    assert ia["t"] == t and ia["dt"] == dt and ia["rejected"] == rej
    if kind != "rk4":
        assert rej > 0  # the retry path (kernel time t + dt) is exercised
    assert np.array_equal(a.get("psi_plus"), b.get("psi_plus"))
    assert np.array_equal(a.get("n_plus"), b.get("n_plus"))
    a.close()
    b.close()


# ----------------------------------------------------------------------------- landing step

def oracle_landing(prec, p, psi, n, P, dt0, out_every, t_max, iterator):
    """main.cu:80-104 with render disabled: iterate until t reaches out_iter * out_every; the
    step that would pass it is shortened to land there and taken with a fixed step (restored dt
    afterwards).  Returns the state and t at every output time."""
    r = rnd(prec)
    # command-line numbers are rounded to the real type (commandline_io.cpp:137); the loop
    # evaluates out_iter * out_every, t + dt and the shortened dt in the real type
    dt0, out_every, t_max = r(dt0), r(out_every), r(t_max)
    t, dt, out_iter = 0.0, dt0, 1
    snaps = []
    while t < t_max:
        target = r(out_iter * out_every)
        while t < target:
            if t >= t_max:
                break
            step_dt, forced = dt, False
            if r(t + dt) > target:
                nd = r(target - t)
                if nd > 0:
                    step_dt, forced = nd, True
            t = r(t + step_dt)
            if iterator == "RK4":
                psi, n = O.rk4_run(p, psi, n, pump=P, dt=step_dt, steps=1)
            elif forced:
                psi, n = O.iter_run(p, "DP45", psi, n, pump=P, dt=step_dt, steps=1)
            else:
                while True:
                    cand = O.dp45_attempt(p, psi, n, pump=P, dt=step_dt)
                    acc, new_dt, _ = O.adapt_dt(cand[2], cand[3], 1e-7, step_dt, 1e-4, 0.2)
                    step_dt = new_dt
                    if acc:
                        psi, n = cand[0], cand[1]
                        break
                dt = step_dt
        out_iter += 1
        snaps.append((t, psi.copy(), n.copy()))
    return snaps


@pytest.mark.parametrize("prec,iterator", [("fp64", "RK4"), ("fp32", "RK4"), ("fp64", "DP45")])
def test_landing_step_matches_oracle(prec, iterator):
    N, L, dt, out_every, t_max = 48, 23.5, 0.04, 0.1, 0.85
    say(f"[landing step] prec={prec} iterator={iterator}: dt={dt} outEvery={out_every}")
    P = pump(N, L)
    psi0 = O.gaussian_field(N, N, 0.01, 17)
    extra = [] if iterator == "RK4" else ["--iterator", "DP45", "-adaptive", "--tol", 1e-7, "--rkvdt", 1e-4, 0.2]
    s = solver(prec, N, L, extra, tmax=t_max, dt=dt, out_every=out_every)
    s.set("psi_plus", psi0)
    s.set("n_plus", np.zeros((N, N)))
    s.set("pump_plus", P)
    got = []
    finished = False
    while not finished:
        _, finished = s.run_interval()
        got.append((s.time()[0], s.get("psi_plus").astype(np.complex128), s.get("n_plus").astype(np.float64)))
    s.close()
    p = O.params(N, N, L, L, pump=True)
    want = oracle_landing(prec, p, psi0, np.zeros((N, N)), P, dt, out_every, t_max, iterator)
    assert len(got) == len(want)
    worst = 0.0
    for (tg, pg, ng), (tw, pw, nw) in zip(got, want):
        assert tg == tw, (tg, tw)
        worst = max(worst, rel_l2(pg, pw), rel_l2(ng, nw))
    say(f"  {len(got)} output times, t = {[round(g[0], 6) for g in got]}; worst rel L2 {worst:.3e}")
    assert worst <= (1e-11 if iterator == "DP45" else TOL[prec])
    # the landing steps happened: 0.1 is not a multiple of 0.04
    assert any(abs(g[0] / out_every - round(g[0] / out_every)) < 1e-6 for g in got[:-1])


# ----------------------------------------------------------------------------- --initRandom

@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_init_random_matches_oracle(prec):
    N, L, dt, steps, amp, seed = 64, 31.5, 0.04, 100, 0.01, 42
    say(f"[initRandom] prec={prec}: --initRandom {amp} {seed}, -tetm")
    s = solver(prec, N, L, ["-tetm", "--initRandom", amp, seed], tmax=(steps - 0.5) * dt)
    gp, gm = s.get("psi_plus").astype(np.complex128), s.get("psi_minus").astype(np.complex128)
    wp, wm = O.init_random(N, N, amp, seed, fp32=prec == "fp32", components=2)
    assert np.array_equal(gp, wp) and np.array_equal(gm, wm)
    P = pump(N, L)
    for k in ("pump_plus", "pump_minus"):
        s.set(k, P)
    assert s.step(steps + 5) == steps
    op, on, om, onm = O.rk4_run(O.params(N, N, L, L, tetm=True, pump=True), wp, np.zeros((N, N)), pump=P, dt=dt,
                                steps=steps, psi_minus=wm, n_minus=np.zeros((N, N)), pump_minus=P)
    e = max(rel_l2(s.get("psi_plus").astype(np.complex128), op), rel_l2(s.get("psi_minus").astype(np.complex128), om))
    s.close()
    say(f"  initial fields bitwise equal; rel L2 after {steps} steps {e:.3e}")
    assert e <= TOL[prec]
