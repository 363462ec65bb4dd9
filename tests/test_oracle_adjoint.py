"""Pins of the FP64 CPU adjoint (orc_rk4_adjoint), CPU only.

The adjoint is checked against what does not depend on it:
  - its padded-grid forward map reproduces orc_rk4_run (the reference-structured
    restatement with subgrids and halos) to 1e-13, periodic and zero boundaries;
  - central finite differences of orc_rk4_run at random pixels, for the pump, the
    potential, Re/Im psi0 and n0 (relative 1e-4 where the FD is well conditioned);
  - linearity: the result is linear in the final-state cotangent;
  - a planted mistake (wrong sign of the potential term in the objective's
    cotangent) is caught by the same FD check.
"""
import numpy as np
import pytest

from oracle import oracle as O

PTH = 0.15 * 0.225 / 0.01


def setup(N, L, periodic, seed=1):
    p = O.params(N, N, L, L, periodic=(periodic, periodic), reservoir=True, pump=True, potential=True)
    rng = np.random.default_rng(seed)
    psi0 = 0.3 * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
    n0 = rng.random((N, N))
    dx = L / (N - 1)
    x = -L / 2 + dx * np.arange(N)
    X, Y = np.meshgrid(x, x)
    P = 2 * PTH * np.exp(-(X ** 2 + Y ** 2) / 20.0)
    V = 3e-4 * np.exp(-((X - 2) ** 2 + (Y + 1) ** 2) / 8.0)
    mask = np.zeros((N, N))
    mask[N // 5:N // 2, N // 3:N - 3] = 1
    return p, psi0, n0, P, V, mask


def objective(p, psi0, n0, P, V, mask, dt, steps):
    a, b = O.rk4_run(p, psi0, n0, pump=P, pot=V, dt=dt, steps=steps)
    return float(np.sum(mask * np.abs(a) ** 2) + 0.3 * np.sum(mask * b))


@pytest.mark.parametrize("periodic", [True, False])
def test_padded_forward_matches_reference_structure(periodic):
    p, psi0, n0, P, V, _ = setup(20, 9.5, periodic)
    o_psi, o_n = O.rk4_run(p, psi0, n0, pump=P, pot=V, dt=0.02, steps=25)
    *_, a_psi, a_n = O.rk4_adjoint(p, psi0, n0, P, V, 0.02, 25, np.zeros_like(psi0))
    assert np.abs(a_psi - o_psi).max() <= 1e-13 * np.abs(o_psi).max()
    assert np.abs(a_n - o_n).max() <= 1e-13 * np.abs(o_n).max()


def fd_check(periodic, plant=False):
    N, L, dt, steps = 20, 9.5, 0.02, 25
    p, psi0, n0, P, V, mask = setup(N, L, periodic)
    psiT, nT = O.rk4_run(p, psi0, n0, pump=P, pot=V, dt=dt, steps=steps)
    lam_psi = 2 * mask * psiT
    if plant:
        lam_psi = np.conj(lam_psi)
    lp, ln, gP, gV, *_ = O.rk4_adjoint(p, psi0, n0, P, V, dt, steps, lam_psi, 0.3 * mask)
    J = lambda **kw: objective(p, kw.get("psi", psi0), kw.get("n", n0), kw.get("P", P), kw.get("V", V), mask, dt, steps)
    rng = np.random.default_rng(7)
    errs = []
    for _ in range(12):
        r, c = rng.integers(0, N, 2)
        checks = []
        for name, base, h, g in (("P", P, 1e-4, gP), ("V", V, 1e-8, gV), ("n", n0, 1e-6, ln)):
            a, b = base.copy(), base.copy()
            a[r, c] += h
            b[r, c] -= h
            checks.append((name, (J(**{name: a}) - J(**{name: b})) / (2 * h), g[r, c]))
        for name, d, g in (("re", 1e-6, lp.real), ("im", 1e-6j, lp.imag)):
            a, b = psi0.copy(), psi0.copy()
            a[r, c] += d
            b[r, c] -= d
            checks.append((name, (J(psi=a) - J(psi=b)) / (2e-6), g[r, c]))
        for name, fd, g in checks:
            scale = {"P": np.abs(gP).max(), "V": np.abs(gV).max(), "n": np.abs(ln).max(), "re": np.abs(lp).max(),
                     "im": np.abs(lp).max()}[name]
            if abs(fd) > 1e-3 * scale:  # well-conditioned central differences only
                errs.append(abs(g - fd) / abs(fd))
    return max(errs), len(errs)


@pytest.mark.parametrize("periodic", [True, False])
def test_adjoint_matches_central_differences(periodic):
    worst, n = fd_check(periodic)
    assert n >= 20
    assert worst <= 1e-4, worst


def test_planted_mistake_is_caught():
    worst, _ = fd_check(True, plant=True)
    assert worst > 1e-2


def test_adjoint_is_linear_in_the_cotangent():
    p, psi0, n0, P, V, mask = setup(16, 7.5, True)
    rng = np.random.default_rng(3)
    l1 = rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))
    l2 = rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))
    a = O.rk4_adjoint(p, psi0, n0, P, V, 0.02, 10, l1)
    b = O.rk4_adjoint(p, psi0, n0, P, V, 0.02, 10, l2)
    c = O.rk4_adjoint(p, psi0, n0, P, V, 0.02, 10, 2.0 * l1 - 3.0 * l2)
    for k in range(4):
        ref = 2.0 * a[k] - 3.0 * b[k]
        assert np.abs(c[k] - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1e-300)
