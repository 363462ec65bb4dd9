"""Pins of the oracle's imaginary-time propagation (--imagTime), CPU only.

With --imagTime the stage sums advance psi by -i dt instead of dt and psi is normalised to
sum |psi|^2 dV = amplitude after every step.  Checked against what the mathematics fixes:
  - two plane waves on a linear periodic problem: every step multiplies mode j by the
    tableau's stability function R(z_j), z_j = -i dt lambda_j (lambda_j the mode's rate), so
    after normalisation the amplitude ratio is (a1/a2) (R(z_1)/R(z_2))^n and the norm is the
    amplitude;
  - a harmonic trap: the iteration converges to the lowest eigenvector of the dense discrete
    Hamiltonian (numpy.linalg.eigh), with the eigenvalue from its Rayleigh quotient;
  - a planted sign error (psi advanced by +i dt) decays the ground-state overlap instead.
"""
import numpy as np
import pytest

from oracle import oracle as O


def stability(key, z):
    """R(z) = 1 + z b^T (I - z A)^-1 1 of the fixed-step tableau, from its Butcher arrays."""
    from test_oracle_iterators import reference_tableau
    A, b, _, _ = reference_tableau(key)
    s = len(b) if key != "BoSha" else 3
    A, b = A[:s, :s], b[:s]
    return 1 + z * (b @ np.linalg.solve(np.eye(s) - z * A, np.ones(s)))


@pytest.mark.parametrize("key", ["RK4", "Heun2", "DP45", "Ralston3"])
def test_two_plane_waves_follow_the_stability_function(key):
    N, L, dt, steps, amp = 16, 7.5, 0.01, 7, 2.5
    p = O.params(N, N, L, L, reservoir=False, g_c=0.0, imag_time_amplitude=amp)
    x = np.arange(N)
    modes = [(1, 0), (3, 2)]
    a = [0.8, 0.6 - 0.3j]
    e = [np.exp(2j * np.pi * (mx * x[None, :] / N + my * x[:, None] / N)) for mx, my in modes]
    psi = a[0] * e[0] + a[1] * e[1]
    out, _ = O.iter_run(p, key, psi, dt=dt, steps=steps)
    assert abs(np.sum(np.abs(out) ** 2) * p.dV - amp) < 1e-12 * amp
    lam = [-1j * p.m_eff_scaled * (-4 * np.sin(np.pi * mx / N) ** 2 / p.dx ** 2 - 4 * np.sin(np.pi * my / N) ** 2 / p.dy ** 2)
           / p.h_bar_s - p.gamma_c / 2 for mx, my in modes]
    R = [stability(key, -1j * dt * l) for l in lam]
    c = [np.vdot(ej, out) / N ** 2 for ej in e]
    expected = a[0] / a[1] * (R[0] / R[1]) ** steps
    assert abs(c[0] / c[1] - expected) < 1e-12 * abs(expected)


def trap_case(N=12):
    L = 5.5
    p = O.params(N, N, L, L, reservoir=False, potential=True, g_c=0.0, gamma_c=0.0, imag_time_amplitude=1.0)
    xs = -L / 2 + p.dx * np.arange(N)
    X, Y = np.meshgrid(xs, xs)
    V = 2e-3 * (X ** 2 + Y ** 2)
    Lap = np.zeros((N * N, N * N))
    for r in range(N):
        for cc in range(N):
            i = r * N + cc
            Lap[i, i] = -2 / p.dx ** 2 - 2 / p.dy ** 2
            for dr, dc, w in ((0, 1, 1 / p.dx ** 2), (0, -1, 1 / p.dx ** 2), (1, 0, 1 / p.dy ** 2), (-1, 0, 1 / p.dy ** 2)):
                rr, c2 = (r + dr) % N, (cc + dc) % N
                Lap[i, rr * N + c2] += w
    H = p.m_eff_scaled * Lap + np.diag(V.ravel())
    return p, V, H


def test_trap_converges_to_the_discrete_ground_state():
    p, V, H = trap_case()
    N = 12
    w, vecs = np.linalg.eigh(H)
    phi0 = vecs[:, 0].reshape(N, N)
    rng = np.random.default_rng(2)
    psi = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    dt = 0.02
    steps = int(np.ceil(30 / (dt * (w[1] - w[0]) / p.h_bar_s)))  # 30 e-folds of the first excited state
    out, _ = O.iter_run(p, "RK4", psi, pot=V, dt=dt, steps=steps)
    overlap = abs(np.vdot(phi0.ravel(), out.ravel())) / np.linalg.norm(out)
    assert 1 - overlap < 1e-10
    E = np.real(np.vdot(out.ravel(), H @ out.ravel())) / np.vdot(out.ravel(), out.ravel()).real
    assert abs(E - w[0]) < 1e-9 * abs(w[1] - w[0])
    assert abs(np.sum(np.abs(out) ** 2) * p.dV - 1.0) < 1e-12


def test_planted_sign_error_moves_away_from_the_ground_state():
    """Advancing by +i dt (the slip the oracle must not make) is a time reversal of the decay:
    the excited content grows, so the overlap with the ground state drops."""
    p, V, H = trap_case()
    N = 12
    w, vecs = np.linalg.eigh(H)
    phi0 = vecs[:, 0].reshape(N, N)
    psi = phi0 + 0.05 * vecs[:, 1].reshape(N, N)
    good, _ = O.iter_run(p, "RK4", psi, pot=V, dt=0.02, steps=200)
    # +i dt: conjugate problem -> run with the negated time step
    bad, _ = O.iter_run(p, "RK4", psi, pot=V, dt=-0.02, steps=200)
    ov = lambda f: abs(np.vdot(phi0.ravel(), f.ravel())) / np.linalg.norm(f)
    assert ov(good) > ov(psi.astype(complex)) > ov(bad)
