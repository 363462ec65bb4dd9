"""Pins of the oracle's complex reservoir (CPU only).

The reference stores the reservoir as Type::complex.  It turns complex when a pump group's temporal
amplitude is complex (kernel_gp_compute.cuh:70-75: k_n += P_g a_g(w) with complex a_g), and in
imaginary time, where runge_sum_to_input_k / runge_add_to_input_k advance n by -i dt like psi
(cuda_macro.cuh:160-178, kernel_summation.cuh:195-258) and normalizeImaginaryTimePropagation rescales
n to the same amplitude (solver_normalize_imag_time.cu:13-45).  With a complex n:
    k_psi = ... + (1/hbar) g_r n (-i psi) + 1/2 R n psi,    k_n = -(gamma_r + R|psi|^2) n + sum_g P_g a_g(w),
and the noise amplitude is the principal square root of (R n + gamma_c)/(4 dV) (kernel_summation.cuh:209-225).
Checks:
  - uniform fields: the textbook RK4 map of the cell ODE in complex arithmetic (numpy), with the stage
    amplitudes a(w) = a0 (1 - w) + w a1 at w = 0, 1/2, 1/2, 1 of RK4;
  - imaginary time: the same map with dt -> -i dt for psi and n, then both normalised;
  - noise: per-cell map with additive stage noise xi * sqrt_c((R n0 + gamma_c)/(4 dV));
  - TE-TM: a uniform spinor with g_pm and two complex reservoirs;
  - split-step iterator: a uniform field follows the closed-form nonlinear map and complex Euler reservoir
    step (gp_scalar_nonlinear with complex n and pump amplitude);
  - real inputs: the complex-reservoir path equals the real-reservoir path (already pinned);
  - planted slips (+i dt on n, the real part of n under the square root, a real pump amplitude) miss.
"""
import numpy as np

from oracle import oracle as O


def rk4_w(f, y, dt, cdt=None):
    """RK4 with stage times w = 0, 1/2, 1/2, 1; cdt is the (complex) step applied to the sums."""
    h = dt if cdt is None else cdt
    k1 = f(y, 0.0)
    k2 = f(y + h / 2 * k1, 0.5)
    k3 = f(y + h / 2 * k2, 0.5)
    k4 = f(y + h * k3, 1.0)
    return y + h / 6 * k1 + h / 3 * k2 + h / 3 * k3 + h / 6 * k4


def scalar_rhs(p, P, a0, a1):
    hb = p.h_bar_s

    def f(y, w):
        psi, n = y
        a = a0 * (1 - w) + w * a1
        dpsi = -1j / hb * (p.g_c * abs(psi) ** 2 + p.g_r * n) * psi + 0.5 * (p.R * n - p.gamma_c) * psi
        dn = -(p.gamma_r + p.R * abs(psi) ** 2) * n + P * a
        return np.array([dpsi, dn])
    return f


def cres_params(N, **kw):
    p = O.params(N, N, 3.5, 3.5, pump=True, **kw)
    p.complex_reservoir = 1
    return p


def test_complex_pump_amplitude_makes_the_reservoir_complex():
    N, dt, steps = 8, 0.04, 3
    p = cres_params(N, g_c=6e-5, g_r=1e-4)
    P = 1.7 * p.gamma_c * p.gamma_r / p.R
    a0, a1 = 0.8 + 0.6j, 0.7 + 0.7j
    O.set_chain(p, "pump", [(np.full((N, N), P), None, a0, a1)])
    y = np.array([2.0 + 1.0j, 3.0 + 0.5j])
    psi, n = O.iter_run(p, "RK4", np.full((N, N), y[0]), np.full((N, N), y[1]), dt=dt, steps=steps)
    assert n.dtype == np.complex128
    f = scalar_rhs(p, P, a0, a1)
    want = y.copy()
    for _ in range(steps):
        want = rk4_w(f, want, dt)
    assert np.abs(psi - want[0]).max() < 1e-12 * abs(want[0])
    assert np.abs(n - want[1]).max() < 1e-12 * abs(want[1])
    assert abs(want[1].imag - 0.5) > 1e-2  # the pump's imaginary part moved Im n
    # planted slip: the pump with the real part of its amplitude only (a real reservoir's rule)
    slip = y.copy()
    fr = scalar_rhs(p, P, a0.real, a1.real)
    for _ in range(steps):
        slip = rk4_w(fr, slip, dt)
    assert abs(slip[1] - want[1]) > 1e-4


def test_imaginary_time_advances_the_reservoir_by_minus_i_dt_and_normalises_it():
    N, dt, steps, amp = 8, 0.04, 4, 3.0
    p = cres_params(N, g_c=6e-5, g_r=1e-4)
    p.imag_time_amplitude = amp
    P = 2.0
    y = np.array([2.0 + 1.0j, 3.0 + 0.5j])
    psi, n = O.iter_run(p, "RK4", np.full((N, N), y[0]), np.full((N, N), y[1]), pump=np.full((N, N), P), dt=dt,
                        steps=steps)
    f = scalar_rhs(p, P, p.pump_amp, p.pump_amp_next)

    def run(cdt_n):
        v = y.copy()
        for _ in range(steps):
            # psi and n share the stage inputs; only n's complex step may differ (the slip below)
            def g(z, w):
                return f(z, w)
            k = [None] * 4
            ws = (0.0, 0.5, 0.5, 1.0)
            steps_c = np.array([-1j * dt, cdt_n])
            z = v.copy()
            for s in range(4):
                k[s] = g(z, ws[s])
                if s < 3:
                    z = v + (0.5 if s < 2 else 1.0) * steps_c * k[s]
            v = v + steps_c * (k[0] / 6 + k[1] / 3 + k[2] / 3 + k[3] / 6)
            v = v * np.sqrt(amp / (N * N * np.abs(v) ** 2 * p.dV))  # each field to the amplitude
        return v
    want = run(-1j * dt)
    assert np.abs(psi - want[0]).max() < 1e-12 * abs(want[0])
    assert np.abs(n - want[1]).max() < 1e-12 * abs(want[1])
    assert abs(np.sum(np.abs(n) ** 2) * p.dV - amp) < 1e-12 * amp
    slip = run(+1j * dt)  # the reservoir advanced by +i dt
    assert abs(slip[1] - want[1]) > 1e-4 * abs(want[1])


def test_complex_reservoir_noise_amplitude_is_the_principal_square_root():
    N, dt, step = 5, 0.04, 11
    # m_eff = 1e30: every cell is its own ODE
    p = cres_params(N, stochastic=0.6, noise_seed=123, m_eff=1e30)
    rng = np.random.default_rng(3)
    psi0 = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    # Re(R n + gamma_c) < 0 in some cells: the square root leaves the positive real axis
    n0 = (rng.standard_normal((N, N)) * 20) + 1j * (rng.standard_normal((N, N)) * 20)
    assert (p.R * n0.real + p.gamma_c < 0).any()
    P = 1 + rng.random((N, N))
    psi, n = O.iter_run(p, "RK4", psi0, n0, pump=P, dt=dt, steps=1, step0=step)
    xi = np.array([O.noise_normal2(p.noise_seed, step, c) for c in range(N * N)]).reshape(N, N) \
        * p.stochastic_amplitude * np.sqrt(dt)
    hb, dV = p.h_bar_s, p.dV

    def per_cell(s):
        inc = xi * s

        def f(y):
            ps, nn = y
            kp = -1j / hb * (p.g_c * abs(ps) ** 2 + p.g_r * nn) * ps + 0.5 * (p.R * nn - p.gamma_c) * ps \
                - p.g_c * (-1j * ps) / (hb * dV)
            kn = -(p.gamma_r + p.R * abs(ps) ** 2) * nn + P * p.pump_amp + p.R * nn / dV
            return np.array([kp, kn])
        y = np.array([psi0, n0])
        k1 = f(y)
        y2 = y + dt / 2 * k1
        y2[0] += 0.5 * inc
        k2 = f(y2)
        y3 = y + dt / 2 * k2
        y3[0] += 0.5 * inc
        k3 = f(y3)
        y4 = y + dt * k3
        y4[0] += inc
        k4 = f(y4)
        out = y + dt / 6 * k1 + dt / 3 * k2 + dt / 3 * k3 + dt / 6 * k4
        out[0] += inc
        return out
    want = per_cell(np.sqrt((p.R * n0 + p.gamma_c) / (4 * dV)))
    assert np.abs(psi - want[0]).max() < 1e-12 * np.abs(want[0]).max()
    assert np.abs(n - want[1]).max() < 1e-12 * np.abs(want[1]).max()
    slip = per_cell(np.sqrt((p.R * n0.real + p.gamma_c + 0j) / (4 * dV)))  # real part of n only
    assert np.abs(slip[0] - want[0]).max() > 1e-6


def test_tetm_uniform_spinor_with_complex_reservoirs_follows_ode():
    N, dt, steps = 8, 0.04, 3
    p = O.params(N, N, 3.5, 4.9, tetm=True, pump=True, g_pm=-2e-5, g_c=6e-5, g_r=1e-4, delta_LT=2.5e-5)
    p.complex_reservoir = 1
    hb = p.h_bar_s
    Pp, Pm = 1.2, 0.7
    a0, a1 = 0.6 - 0.8j, 0.5 - 0.9j
    O.set_chain(p, "pump", [(np.full((N, N), Pp), np.full((N, N), Pm), a0, a1)])
    c0 = 2 / p.dx ** 2 - 2 / p.dy ** 2

    def f(y, w):
        sp, sm, npl, nmi = y
        a = a0 * (1 - w) + w * a1
        dsp = -1j / hb * (p.g_c * abs(sp) ** 2 + p.g_r * npl + p.g_pm * abs(sm) ** 2) * sp \
            + 0.5 * (p.R * npl - p.gamma_c) * sp - 1j / hb * p.delta_LT * c0 * sm
        dsm = -1j / hb * (p.g_c * abs(sm) ** 2 + p.g_r * nmi + p.g_pm * abs(sp) ** 2) * sm \
            + 0.5 * (p.R * nmi - p.gamma_c) * sm - 1j / hb * p.delta_LT * c0 * sp
        dnp = -(p.gamma_r + p.R * abs(sp) ** 2) * npl + Pp * a
        dnm = -(p.gamma_r + p.R * abs(sm) ** 2) * nmi + Pm * a
        return np.array([dsp, dsm, dnp, dnm])

    y = np.array([3.0 + 1.0j, -1.0 + 2.0j, 10.0 - 1.0j, 7.0 + 2.0j])
    full = lambda v: np.full((N, N), v)
    op, onp, om, onm = O.iter_run(p, "RK4", full(y[0]), full(y[2]), dt=dt, steps=steps, psi_minus=full(y[1]),
                                  n_minus=full(y[3]))
    for _ in range(steps):
        y = rk4_w(f, y, dt)
    for got, want in ((op, y[0]), (om, y[1]), (onp, y[2]), (onm, y[3])):
        assert np.abs(got - want).max() < 1e-12 * max(1.0, abs(want)), (got.flat[0], want)


def test_real_inputs_give_the_real_reservoir_result():
    """Real amplitudes and a real n0: the complex-reservoir path keeps Im n = 0 and equals the real path."""
    N, dt, steps = 12, 0.04, 6
    rng = np.random.default_rng(1)
    psi0 = 0.5 * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
    n0 = 3 + rng.random((N, N))
    P = 1 + rng.random((N, N))
    out = {}
    for cres in (0, 1):
        p = O.params(N, N, 5.5, 5.5, pump=True, periodic=(False, True))
        p.complex_reservoir = cres
        out[cres] = O.iter_run(p, "DP45", psi0, n0, pump=P, dt=dt, steps=steps)
    assert np.abs(out[1][1].imag).max() == 0.0
    assert np.abs(out[1][0] - out[0][0]).max() <= 1e-14 * np.abs(out[0][0]).max()
    assert np.abs(out[1][1].real - out[0][1]).max() <= 1e-14 * np.abs(out[0][1]).max()


def test_split_step_uniform_field_with_complex_pump_amplitude():
    """Split-step iterator (gp_scalar_nonlinear, kernel_gp_compute.cuh:264-305) with a complex reservoir: a
    uniform field is untouched by the linear half steps, so one step is
        psi <- psi exp(dt (-(i / hbar) (g_c |psi|^2 + g_r n) + (R n - gamma_c) / 2)),
        n <- n + dt (-(gamma_r + R |psi|^2) n + P a)
    in complex arithmetic, a the pump amplitude of the step start."""
    N, L, dt, steps = 16, 7.5, 0.03, 8
    p = O.params(N, N, L, L, pump=True, g_c=6e-5, g_r=1e-4)
    p.complex_reservoir = 1
    P, a = 3.0, 0.6 + 0.8j
    O.set_chain(p, "pump", [(np.full((N, N), P), None, a, a)])
    psi0, n0 = 0.4 - 0.2j, 0.8 + 0.3j
    out, n = O.ssfm_run(p, np.full((N, N), psi0), np.full((N, N), n0), dt=dt, steps=steps)
    assert n.dtype == np.complex128

    def run(amp):
        x, y = psi0, n0
        for _ in range(steps):
            rho = abs(x) ** 2
            x_new = x * np.exp(dt * (-1j / p.h_bar_s * (p.g_c * rho + p.g_r * y) + (p.R * y - p.gamma_c) / 2))
            y = y + dt * (-(p.gamma_r + p.R * rho) * y + P * amp)
            x = x_new
        return x, y
    x, y = run(a)
    assert np.abs(out - x).max() < 1e-13 * abs(x)
    assert np.abs(n - y).max() < 1e-13 * abs(y)
    slip = run(a.real)  # the pump's real part only (a real reservoir's rule)
    assert abs(slip[1] - y) > 1e-3 * abs(y) and abs(slip[0] - x) > 1e-6 * abs(x)
