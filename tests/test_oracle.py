"""Pins of the FP64 oracle against what the mathematics fixes (CPU only).

Each test is chosen so that a plausible slip in the oracle (a dropped term, a
wrong sign, a wrong stage extent or weight, a wrong boundary) fails it:
  - Philox4x32-10 known-answer vectors (Salmon et al. / Random123 kat_vectors);
  - RK4 on a linear periodic problem = the degree-4 Taylor polynomial of the
    stencil matrix (a textbook identity, evaluated with numpy's dense algebra);
  - the same with zero boundaries: the reference's evolving halo ring makes one
    step equal to RK4 on the infinitely zero-padded lattice, cropped;
  - plane waves decay by the RK4 stability polynomial R(z);
  - the homogeneous condensate fixed point n* = gamma_c/R, |psi|^2 = P/gamma_c - gamma_r/R
    rotates with omega = (g_c|psi|^2 + g_r n*)/hbar;
  - the reservoir-only ODE has the RK4 closed form;
  - results are independent of the subgrid decomposition (bitwise);
  - the FFT filter equals numpy's ifft2(fft2(psi) * ifftshift(mask));
  - Dormand-Prince 5(4): the 5th-order stability polynomial and the order of the
    embedded error estimate; the step-size controller law.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def test_philox_known_answers():
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_noise_is_standard_complex_normal():
    z = np.array([O.noise_normal2(7, 3, c) for c in range(20000)])
    for part in (z.real, z.imag):
        assert abs(part.mean()) < 0.03 and abs(part.std() - 1) < 0.03
    assert abs(np.mean(z.real * z.imag)) < 0.03


def laplacian_matrix(Nc, Nr, dx, dy, periodic, pad=0):
    """Dense 5-point Laplacian on an (Nr+2pad) x (Nc+2pad) lattice, row-major."""
    W, H = Nc + 2 * pad, Nr + 2 * pad
    n = W * H
    A = np.zeros((n, n))
    for r in range(H):
        for c in range(W):
            i = r * W + c
            A[i, i] = -2 / dx ** 2 - 2 / dy ** 2
            for dr, dc, w in ((0, 1, 1 / dx ** 2), (0, -1, 1 / dx ** 2), (1, 0, 1 / dy ** 2), (-1, 0, 1 / dy ** 2)):
                rr, cc = r + dr, c + dc
                if periodic:
                    rr %= H
                    cc %= W
                elif not (0 <= rr < H and 0 <= cc < W):
                    continue
                A[i, rr * W + cc] += w
    return A


def taylor4(M):
    I = np.eye(M.shape[0])
    M2 = M @ M
    return I + M + M2 / 2 + M2 @ M / 6 + M2 @ M2 / 24


@pytest.mark.parametrize("periodic", [True, False])
def test_rk4_linear_equals_taylor_polynomial(periodic):
    Nc, Nr, L, dt = 10, 8, 4.5, 0.05
    p = O.params(Nc, Nr, L, L * (Nr - 1) / (Nc - 1), periodic=(periodic, periodic), reservoir=False, potential=True,
                 g_c=0.0)
    rng = np.random.default_rng(3)
    psi = rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc))
    V = 2e-4 * rng.random((Nr, Nc))
    out, _ = O.rk4_run(p, psi, pot=V, dt=dt, steps=1)
    pad = 0 if periodic else 4
    Lap = laplacian_matrix(Nc, Nr, p.dx, p.dy, periodic, pad)
    Vp = np.pad(V, pad)
    # k = -i/hbar (m~ Lap + V) psi - gamma_c/2 psi
    A = -1j / p.h_bar_s * (p.m_eff_scaled * Lap + np.diag(Vp.ravel())) - p.gamma_c / 2 * np.eye(Lap.shape[0])
    ref = (taylor4(dt * A) @ np.pad(psi, pad).ravel()).reshape(Nr + 2 * pad, Nc + 2 * pad)
    if pad:
        ref = ref[pad:-pad, pad:-pad]
    assert np.abs(out - ref).max() < 1e-12 * np.abs(ref).max()


def test_zero_boundary_differs_from_dirichlet_ghost():
    """The reference's zero-BC halo evolves within a step: not the same as a fixed zero ghost layer."""
    Nc = Nr = 8
    p = O.params(Nc, Nr, 3.5, 3.5, periodic=(False, False), reservoir=False, g_c=0.0)
    psi = np.ones((Nr, Nc), complex)
    out, _ = O.rk4_run(p, psi, dt=0.05, steps=1)
    A = -1j / p.h_bar_s * p.m_eff_scaled * laplacian_matrix(Nc, Nr, p.dx, p.dy, False, 0) - p.gamma_c / 2 * np.eye(64)
    ghost = (taylor4(0.05 * A) @ psi.ravel()).reshape(Nr, Nc)
    assert np.abs(out - ghost).max() > 1e-6


def test_plane_wave_decays_by_stability_polynomial():
    N, L, dt, steps = 16, 7.5, 0.01, 3
    p = O.params(N, N, L, L, reservoir=False, g_c=0.0)
    x = np.arange(N)
    mx, my = 3, 2
    psi = np.exp(2j * np.pi * (mx * x[None, :] / N + my * x[:, None] / N))
    out, _ = O.rk4_run(p, psi, dt=dt, steps=steps)
    lam = -4 * np.sin(np.pi * mx / N) ** 2 / p.dx ** 2 - 4 * np.sin(np.pi * my / N) ** 2 / p.dy ** 2
    z = dt * (-1j * p.m_eff_scaled * lam / p.h_bar_s - p.gamma_c / 2)
    R = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    assert np.abs(out - psi * R ** steps).max() < 1e-13


def test_homogeneous_condensate_fixed_point():
    N, L, dt, steps = 8, 3.5, 0.04, 200
    p = O.params(N, N, L, L, pump=True)
    Pth = p.gamma_c * p.gamma_r / p.R
    P0 = 1.5 * Pth
    nstar = p.gamma_c / p.R
    rho = P0 / p.gamma_c - p.gamma_r / p.R
    psi = np.full((N, N), np.sqrt(rho) + 0j)
    out, n = O.rk4_run(p, psi, np.full((N, N), nstar), pump=np.full((N, N), P0), dt=dt, steps=steps)
    omega = (p.g_c * rho + p.g_r * nstar) / p.h_bar_s
    assert np.abs(n - nstar).max() < 1e-8
    assert np.abs(out - np.sqrt(rho) * np.exp(-1j * omega * dt * steps)).max() < 1e-9


def test_reservoir_ode_closed_form():
    N, dt, steps = 6, 0.1, 7
    p = O.params(N, N, 2.5, 2.5, pump=True)
    P = np.linspace(0.5, 2.0, N * N).reshape(N, N)
    n0 = np.linspace(0.0, 3.0, N * N).reshape(N, N)
    _, n = O.rk4_run(p, np.zeros((N, N), complex), n0, pump=P, dt=dt, steps=steps)
    z = -p.gamma_r * dt
    R = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    S = dt * (1 + z / 2 + z * z / 6 + z ** 3 / 24)
    ref = n0.copy()
    for _ in range(steps):
        ref = R * ref + S * P
    assert np.abs(n - ref).max() < 1e-13


@pytest.mark.parametrize("periodic", [True, False])
def test_subgrid_decomposition_invariance(periodic):
    N, L = 32, 15.5
    rng = np.random.default_rng(5)
    psi = 0.1 * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
    n = rng.random((N, N))
    P = 5 * rng.random((N, N))
    V = 1e-4 * rng.random((N, N))
    outs = []
    for sg in [(1, 1), (4, 4), (2, 4), (8, 2)]:
        p = O.params(N, N, L, L, periodic=(periodic, periodic), subgrids=sg, pump=True, potential=True)
        outs.append(O.rk4_run(p, psi, n, pump=P, pot=V, dt=0.04, steps=5))
    for o in outs[1:]:
        assert np.array_equal(o[0], outs[0][0]) and np.array_equal(o[1], outs[0][1])


def test_translation_equivariance_periodic():
    N, L = 24, 11.5
    rng = np.random.default_rng(6)
    psi = 0.1 * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
    n, P = rng.random((N, N)), 5 * rng.random((N, N))
    p = O.params(N, N, L, L, pump=True)
    a, an = O.rk4_run(p, psi, n, pump=P, dt=0.04, steps=4)
    sh = (5, 11)
    b, bn = O.rk4_run(p, np.roll(psi, sh, (0, 1)), np.roll(n, sh, (0, 1)), pump=np.roll(P, sh, (0, 1)), dt=0.04, steps=4)
    assert np.array_equal(np.roll(a, sh, (0, 1)), b) and np.array_equal(np.roll(an, sh, (0, 1)), bn)


@pytest.mark.parametrize("shape", [(16, 16), (32, 8), (12, 10)])
def test_fft2d_matches_numpy(shape):
    rng = np.random.default_rng(1)
    z = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    assert np.allclose(O.fft2d(z, -1), np.fft.fft2(z), atol=1e-10)
    assert np.allclose(O.fft2d(z, +1), np.fft.ifft2(z) * z.size, atol=1e-10)


def test_fft_filter_matches_numpy():
    rng = np.random.default_rng(2)
    Nr, Nc = 16, 32
    psi = rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc))
    mask = rng.random((Nr, Nc))
    out = O.fft_filter(psi, mask)
    ref = np.fft.ifft2(np.fft.fft2(psi) * np.fft.ifftshift(mask))
    assert np.allclose(out, ref, atol=1e-12)


def test_envelope_gaussian_and_ring():
    N, L = 33, 16.0
    dx = L / (N - 1)
    x = -L / 2 + dx * np.arange(N)
    X, Y = np.meshgrid(x, x)
    g = O.envelope(N, N, L, L, dx, dx, 2.0, 3.0, 2.0, 1.0, -0.5, 1.0, O.ENV_GAUSS | O.ENV_NODIVIDE)
    f = 0.5 * (((X - 1.0) / 3.0) ** 2 + ((Y + 0.5) / 2.0) ** 2)
    assert np.allclose(g, 2.0 * np.exp(-f), atol=1e-14)
    r = O.envelope(N, N, L, L, dx, dx, 2.0, 3.0, 3.0, 0, 0, 2.0, O.ENV_GAUSS | O.ENV_RING, charge=1.0)
    f = 0.5 * ((X / 3.0) ** 2 + (Y / 3.0) ** 2)
    ref = 2.0 / np.sqrt(2 * 3.1415 * 9.0) * f * np.exp(-f ** 2) * np.exp(1j * np.arctan2(X, Y))
    assert np.allclose(r, ref, atol=1e-14)


# ---------------------------------------------------------------------------- DP45

def dp45_scalar_response(z):
    """One adaptive-DP45 attempt on the 1-cell linear ODE psi' = (-i a - gamma_c/2) psi, driven through the oracle."""
    p = O.params(8, 8, 3.5, 3.5, reservoir=False, g_c=0.0, potential=True)
    # on a uniform periodic field the Laplacian vanishes; V sets the rotation rate
    gamma = p.gamma_c
    dt = 0.1
    lam = z / dt  # desired eigenvalue -i V/hbar - gamma/2  ->  V = (-(lam + gamma/2)/(-1j)) * hbar
    V = ((lam + gamma / 2) / (-1j)).real * p.h_bar_s
    psi = np.ones((8, 8), complex)
    out, _, err, norm = O.dp45_attempt(p, psi, None, pot=np.full((8, 8), V), dt=dt)
    return out[0, 0], err / 64, dt * (-1j * V / p.h_bar_s - gamma / 2)


def test_dp45_fifth_order_stability_polynomial():
    for y in (0.05, 0.2, 0.5):
        out, _, z = dp45_scalar_response(-1j * y)
        R5 = sum(z ** k / math.factorial(k) for k in range(6)) + z ** 6 / 600
        assert abs(out - R5) < 1e-14


def test_dp45_error_estimate_is_fifth_order():
    # err = |R5(z) - R4*(z)|^2 with R5 - R4* = O(z^5): halving z divides sqrt(err) by ~32
    e = []
    for y in (0.2, 0.1, 0.05):
        _, err, _ = dp45_scalar_response(-1j * y)
        e.append(np.sqrt(err))
    assert 28 < e[0] / e[1] < 36 and 29 < e[1] / e[2] < 35


def test_adapt_dt_controller_law():
    acc, nd, fe = O.adapt_dt(err_sum=1e-8, norm=1.0, tol=1e-6, dt=0.01, dt_min=1e-4, dt_max=0.1)
    assert acc and abs(fe - 1e-8) < 1e-20 and abs(nd - 0.01 * (1e-6 / 2 / 1e-8) ** 0.25) < 1e-15
    acc, nd, _ = O.adapt_dt(err_sum=1e-2, norm=1.0, tol=1e-6, dt=0.04, dt_min=1e-4, dt_max=0.1)
    assert not acc and nd < 0.04
    acc, nd, fe = O.adapt_dt(err_sum=float("nan"), norm=1.0, tol=1e-6, dt=0.04, dt_min=1e-4, dt_max=0.1)
    assert not acc and abs(nd - 0.036) < 1e-15  # NaN -> dh = 0.9, reject (solver.cu:81-88)
    acc, nd, _ = O.adapt_dt(err_sum=1e-30, norm=1.0, tol=1e-6, dt=0.04, dt_min=1e-4, dt_max=0.1)
    assert acc and nd == 0.1  # clamped to dt_max
