"""Pins of the oracle's split-step Fourier iterator (orc_ssfm_run), CPU only.

  - a free plane wave exp(2 pi i m col / N_c) acquires exp(-i hbar k^2 / (2 m) dt) per step,
    k = 2 * 3.1415926535 * m / L_x (the constant the reference writes);
  - on a uniform field only the k = 0 mode is present, the linear half steps are the identity
    and a step is the closed-form nonlinear map psi exp(dt (R n - gamma_c)/2) exp(-i dt A / hbar)
    with the explicit Euler reservoir update;
  - a linear problem with a potential equals the textbook Strang splitting computed with
    numpy.fft (independent transform, the same ordering: half kinetic, full potential, half
    kinetic), on a non-square grid;
  - the noise increment is xi sqrt((R n + gamma_c) / (4 dV)) with the step-start reservoir;
  - with -tetm the reservoirs keep their input values (the reference's quirk).
"""
import numpy as np
import pytest

from oracle import oracle as O


def test_free_plane_wave_phase():
    Nc, Nr, L, dt, steps = 32, 16, 15.5, 0.05, 9
    Ly = L * (Nr - 1) / (Nc - 1)
    p = O.params(Nc, Nr, L, Ly, reservoir=False, g_c=0.0, gamma_c=0.0)
    m = 3
    col = np.arange(Nc)
    psi = np.tile(np.exp(2j * np.pi * m * col / Nc), (Nr, 1))
    out, _ = O.ssfm_run(p, psi, dt=dt, steps=steps)
    k = 2 * 3.1415926535 * m / L
    w = p.h_bar_s / 2 / p.m_eff * k * k
    assert np.abs(out - psi * np.exp(-1j * w * dt * steps)).max() < 1e-12


def test_uniform_field_closed_form():
    N, L, dt, steps = 16, 7.5, 0.03, 12
    p = O.params(N, N, L, L, pump=True, potential=True, pump_amp=1.3, pot_amp=0.7)
    psi0, n0, P, V = 0.4 - 0.2j, 0.8, 3.0, 2e-3
    out, n = O.ssfm_run(p, np.full((N, N), psi0), np.full((N, N), n0), pump=np.full((N, N), P), pot=np.full((N, N), V),
                        dt=dt, steps=steps)
    x, y = psi0, n0
    for _ in range(steps):
        rho = abs(x) ** 2
        A = p.g_c * rho + V * 0.7 + p.g_r * y
        x_new = x * np.exp(dt * (p.R * y - p.gamma_c) / 2) * np.exp(-1j * dt * A / p.h_bar_s)
        y = y + dt * (-(p.gamma_r + p.R * rho) * y + P * 1.3)
        x = x_new
    assert np.abs(out - x).max() < 1e-13 * abs(x)
    assert np.abs(n - y).max() < 1e-13 * abs(y)


def test_uniform_field_imaginary_time_is_the_real_map_then_normalisation():
    """--imagTime with the split-step iterator: the step takes no -i dt (split_step_fourier_method.cu:25-71),
    then psi and the (real) reservoir are each rescaled to sum |f|^2 dV = amp (solver_normalize_imag_time.cu)."""
    N, L, dt, steps, amp = 16, 7.5, 0.03, 6, 2.5
    p = O.params(N, N, L, L, pump=True, pump_amp=1.3)
    p.imag_time_amplitude = amp
    psi0, n0, P = 0.4 - 0.2j, 0.8, 3.0
    out, n = O.ssfm_run(p, np.full((N, N), psi0), np.full((N, N), n0), pump=np.full((N, N), P), dt=dt, steps=steps)
    assert n.dtype == np.float64

    def run(normalise_n):
        x, y = psi0, n0
        for _ in range(steps):
            rho = abs(x) ** 2
            A = p.g_c * rho + p.g_r * y
            x_new = x * np.exp(dt * (p.R * y - p.gamma_c) / 2) * np.exp(-1j * dt * A / p.h_bar_s)
            y = y + dt * (-(p.gamma_r + p.R * rho) * y + P * 1.3)
            x = x_new * np.sqrt(amp / (N * N * abs(x_new) ** 2 * p.dV))
            if normalise_n:
                y = y * np.sqrt(amp / (N * N * y ** 2 * p.dV))
        return x, y
    x, y = run(True)
    assert np.abs(out - x).max() < 1e-13 * abs(x)
    assert np.abs(n - y).max() < 1e-13 * abs(y)
    assert abs(run(False)[1] - y) > 1e-3 * abs(y)  # the reservoir's normalisation matters


def test_linear_problem_is_strang_splitting():
    Nc, Nr, L, dt, steps = 32, 16, 15.5, 0.04, 5
    Ly = L * (Nr - 1) / (Nc - 1)
    p = O.params(Nc, Nr, L, Ly, reservoir=False, potential=True, g_c=0.0)
    rng = np.random.default_rng(4)
    psi = rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc))
    xs = -L / 2 + p.dx * np.arange(Nc)
    ys = -Ly / 2 + p.dy * np.arange(Nr)
    X, Y = np.meshgrid(xs, ys)
    V = 1e-3 * (X ** 2 + Y ** 2) / 10
    out, _ = O.ssfm_run(p, psi, pot=V, dt=dt, steps=steps)
    kx = 2 * 3.1415926535 * np.fft.fftfreq(Nc, 1.0 / Nc) / L
    ky = 2 * 3.1415926535 * np.fft.fftfreq(Nr, 1.0 / Nr) / Ly
    # numpy's fftfreq puts N/2 at -N/2; the reference maps index N/2 to +N/2: same k^2
    K2 = kx[None, :] ** 2 + ky[:, None] ** 2
    half = np.exp(-1j * p.h_bar_s / 2 / p.m_eff * K2 * dt / 2)
    ref = psi.copy()
    for _ in range(steps):
        ref = np.fft.ifft2(np.fft.fft2(ref) * half)
        ref = ref * np.exp(-p.gamma_c * dt / 2) * np.exp(-1j * V * dt / p.h_bar_s)
        ref = np.fft.ifft2(np.fft.fft2(ref) * half)
    assert np.abs(out - ref).max() < 1e-12 * np.abs(ref).max()


def test_noise_increment_uses_the_step_start_reservoir():
    N, L, dt, seed, amp = 8, 3.5, 0.02, 5, 0.3
    p = O.params(N, N, L, L, pump=True, stochastic=amp, noise_seed=seed)
    p0 = O.params(N, N, L, L, pump=True)
    rng = np.random.default_rng(1)
    psi = 0.1 * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
    n = rng.random((N, N))
    P = rng.random((N, N))
    a, an = O.ssfm_run(p, psi, n, pump=P, dt=dt, steps=1, step0=4)
    # the same step without the noise increment (but with the noise's drift corrections)
    pd = O.params(N, N, L, L, pump=True, stochastic=amp, noise_seed=seed)
    xi = np.array([O.noise_normal2(seed, 4, c) for c in range(N * N)]).reshape(N, N) * amp * np.sqrt(dt)
    sg = np.sqrt((p.R * n + p.gamma_c) / (4 * p.dV))
    # reconstruct: a - xi*sg must be the deterministic part, which does not depend on the seed
    pd.noise_seed = seed + 1
    b, bn = O.ssfm_run(pd, psi, n, pump=P, dt=dt, steps=1, step0=4)
    xi_b = np.array([O.noise_normal2(seed + 1, 4, c) for c in range(N * N)]).reshape(N, N) * amp * np.sqrt(dt)
    assert np.abs((a - xi * sg) - (b - xi_b * sg)).max() < 1e-14
    assert np.array_equal(an, bn)
    assert np.abs(xi * sg).max() > 1e-3


def test_tetm_reservoir_is_not_installed():
    N, L, dt = 8, 3.5, 0.02
    p = O.params(N, N, L, L, pump=True, tetm=True)
    rng = np.random.default_rng(2)
    psi, psim = 0.1 * rng.standard_normal((N, N)) + 0j, 0.1 * rng.standard_normal((N, N)) + 0j
    n, nm, P = rng.random((N, N)), rng.random((N, N)), rng.random((N, N))
    _, n1, _, nm1 = O.ssfm_run(p, psi, n, pump=P, dt=dt, steps=3, psi_minus=psim, n_minus=nm, pump_minus=P)
    assert np.array_equal(n1, n) and np.array_equal(nm1, nm)
    ps = O.params(N, N, L, L, pump=True)
    _, n2 = O.ssfm_run(ps, psi, n, pump=P, dt=dt, steps=3)
    assert np.abs(n2 - n).max() > 1e-4


def test_uniform_field_source_chains_and_pulse_closed_form():
    """Chains of temporal groups at the amplitude of the step start (gp_scalar_nonlinear /
    gp_scalar_independent, kernel_gp_compute.cuh:264-440): the potential enters as V_g a_g with complex
    a_g (its imaginary part is a gain dt Im(a) V / hbar), the pump as P_g Re a_g, the pulse as
    (dt / hbar) F_g a_g in the scalar model."""
    N, L, dt, steps = 16, 7.5, 0.03, 3
    p = O.params(N, N, L, L, pump=True, potential=True)
    psi0, n0 = 0.4 - 0.2j, 0.8
    Vg, av = [2e-3, -1e-3], [0.7 + 0.2j, 1.1 - 0.4j]
    Pg, ap = [3.0, 1.5], [1.3, 0.6]
    Fg, al = [0.02 + 0.01j, -0.03j], [0.5 - 0.5j, 0.9 + 0.1j]
    full = lambda v: np.full((N, N), v)
    O.set_chain(p, "potential", [(full(v), None, a, a) for v, a in zip(Vg, av)])
    O.set_chain(p, "pump", [(full(v), None, a, a) for v, a in zip(Pg, ap)])
    O.set_chain(p, "pulse", [(full(v), None, a, a) for v, a in zip(Fg, al)])
    out, n = O.ssfm_run(p, full(psi0), full(n0), dt=dt, steps=steps)
    x, y = psi0, n0
    for _ in range(steps):
        rho = abs(x) ** 2
        A = p.g_c * rho + sum(v * a.real for v, a in zip(Vg, av)) + p.g_r * y
        B = sum(v * a.imag for v, a in zip(Vg, av))
        x_new = x * np.exp(dt * (p.R * y - p.gamma_c) / 2 + dt * B / p.h_bar_s) * np.exp(-1j * dt * A / p.h_bar_s)
        x_new += dt / p.h_bar_s * sum(f * a for f, a in zip(Fg, al))
        y = y + dt * (-(p.gamma_r + p.R * rho) * y + sum(v * a for v, a in zip(Pg, ap)))
        x = x_new
    assert np.abs(out - x).max() < 1e-13 * abs(x)
    assert np.abs(n - y).max() < 1e-13 * abs(y)


def test_tetm_pulse_enters_times_minus_i():
    """The TE-TM independent step adds -i (dt / hbar) F_g a_g to each component (kernel_gp_compute.cuh:409-413)."""
    N, L, dt = 16, 7.5, 0.03
    p = O.params(N, N, L, L, tetm=True, reservoir=False, g_c=0.0, gamma_c=0.0)
    p.g_pm = 0.0
    Fp, Fm, a = 0.02 + 0.01j, -0.03j, 0.5 - 0.5j
    full = lambda v: np.full((N, N), v)
    O.set_chain(p, "pulse", [(full(Fp), full(Fm), a, a)])
    out_p, _, out_m, _ = O.ssfm_run(p, full(0.1 + 0j), psi_minus=full(0.2j), dt=dt)
    assert np.abs(out_p - (0.1 - 1j * dt / p.h_bar_s * Fp * a)).max() < 1e-15
    assert np.abs(out_m - (0.2j - 1j * dt / p.h_bar_s * Fm * a)).max() < 1e-15
