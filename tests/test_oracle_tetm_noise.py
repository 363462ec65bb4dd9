"""Pins of the oracle's TE-TM right-hand side and of its stochastic terms (CPU only).

TE-TM branch (kernel_gp_compute.cuh:112-235): for psi+- the RHS is
    k+- = -(i/hbar)[m~ Lap psi+- + (g_c|psi+-|^2 + g_r n+- + g_pm|psi-+|^2 + V+-) psi+-]
          + (R n+- - gamma_c)/2 psi+- - (i/hbar) delta_LT cross+-,
    cross+- = (hor - ver)(psi-+) +- (i/2)/(dx dy) (psi-+[r+1,c+1] + psi-+[r-1,c-1] - psi-+[r+1,c-1] - psi-+[r-1,c+1]),
hor/ver the E+W / N+S neighbour sums over dx^2 / dy^2 without a centre term.
  - A two-component plane wave exp(i(kx x + ky y)) is an eigenvector of the discrete operators:
    Lap -> -4 sin^2(a/2)/dx^2 - 4 sin^2(b/2)/dy^2, hor - ver -> 2cos a/dx^2 - 2cos b/dy^2 and the diagonal
    combination -> -4 sin a sin b (a = kx dx, b = ky dy).  With the nonlinear terms off, one RK4 step is
    then the degree-4 Taylor polynomial of dt times the 2x2 symbol, a closed form independent of any
    stencil code (dx != dy, both wave-vector signs).
  - With zero boundaries, one step equals the Taylor polynomial of the dense 2N x 2N operator on the
    lattice padded by the halo (the reference's evolving halo ring, diagonal corners included).
  - A uniform spinor with g_pm, reservoirs, an elliptic pump and potentials follows the textbook RK4 map
    of its 4-variable ODE; the uniform field's cross term (2/dx^2 - 2/dy^2) psi-+ is the reference's
    missing centre term when dx != dy.
  - A planted slip (the sign of the (i/2) term) and a flipped kinetic sign are far outside the tolerance.
Stochastic terms (kernel_summation.cuh:209-225, kernel_gp_compute.cuh:102-108,176-181):
  - with the kinetic term negligible (huge m_eff) every cell is its own ODE; one step equals the
    textbook RK4 map whose stage inputs and result each add (sum of the stage weights) * xi * s,
    s = sqrt((R n0 + gamma_c)/(4 dV)) with the step-start reservoir, xi = amp sqrt(dt) N2(seed, step, cell)
    from the Philox stream pinned by its known answers, and whose RHS carries the drifts
    -g_c(-i psi)/(hbar dV) and +R n/dV;
  - with -tetm both components get the same xi and the psi+ reservoir's amplitude (reference quirk 6).
"""
import numpy as np
import pytest

from oracle import oracle as O


def taylor4(M):
    I = np.eye(M.shape[0], dtype=complex)
    M2 = M @ M
    return I + M + M2 / 2 + M2 @ M / 6 + M2 @ M2 / 24


def rk4_map(f, y, dt):
    k1 = f(y)
    k2 = f(y + dt / 2 * k1)
    k3 = f(y + dt / 2 * k2)
    k4 = f(y + dt * k3)
    return y + dt / 6 * k1 + dt / 3 * k2 + dt / 3 * k3 + dt / 6 * k4


# ----------------------------------------------------------------------------- TE-TM plane waves

def tetm_symbol(p, a, b, V0, i_half_sign=1.0, kin_sign=1.0):
    """2x2 symbol of the linear TE-TM RHS for the plane wave with phase steps a = kx dx, b = ky dy."""
    hb = p.h_bar_s
    lam = -4 * np.sin(a / 2) ** 2 / p.dx ** 2 - 4 * np.sin(b / 2) ** 2 / p.dy ** 2
    hx, hy = 2 * np.cos(a) / p.dx ** 2, 2 * np.cos(b) / p.dy ** 2
    mixed = 2 * np.sin(a) * np.sin(b) / (p.dx * p.dy)  # (1/2)/(dx dy) * 4 sin a sin b
    d = -1j / hb * (kin_sign * p.m_eff_scaled * lam + V0) - p.gamma_c / 2
    op = -1j / hb * p.delta_LT * (hx - hy - i_half_sign * 1j * mixed)  # psi+ row: couples psi-
    om = -1j / hb * p.delta_LT * (hx - hy + i_half_sign * 1j * mixed)  # psi- row: couples psi+
    return np.array([[d, op], [om, d]])


def plane_wave_case(mx, my, dt=0.04, delta_LT=2e-2):
    Nc, Nr = 16, 12
    Lx, Ly = 7.5, 6.6  # dx = 0.5, dy = 0.6
    p = O.params(Nc, Nr, Lx, Ly, tetm=True, reservoir=True, potential=True, g_c=0.0, g_pm=0.0, g_r=0.0, R=0.0,
                 delta_LT=delta_LT)
    c = np.arange(Nc)[None, :]
    r = np.arange(Nr)[:, None]
    wave = np.exp(2j * np.pi * (mx * c / Nc + my * r / Nr))
    ap, am = 0.7 - 0.2j, -0.3 + 0.5j
    V0 = 3e-4
    V = np.full((Nr, Nc), V0)
    a, b = 2 * np.pi * mx / Nc, 2 * np.pi * my / Nr
    return p, wave, ap, am, V, V0, a, b, dt


@pytest.mark.parametrize("mx,my", [(3, 2), (-3, 2), (5, -1), (2, 5)])
def test_tetm_plane_wave_equals_symbol_taylor_polynomial(mx, my):
    p, wave, ap, am, V, V0, a, b, dt = plane_wave_case(mx, my)
    n0 = np.zeros(wave.shape)
    op, _, om, _ = O.rk4_run(p, ap * wave, n0, pot=V, pot_minus=V, dt=dt, steps=1, psi_minus=am * wave, n_minus=n0)
    T = taylor4(dt * tetm_symbol(p, a, b, V0))
    cp, cm = T @ np.array([ap, am])
    scale = abs(cp) + abs(cm)
    assert np.abs(op - cp * wave).max() < 1e-13 * scale
    assert np.abs(om - cm * wave).max() < 1e-13 * scale
    # the coupling is visible: without delta_LT the components would not mix by this much
    T0 = taylor4(dt * tetm_symbol(O.params(16, 12, 7.5, 6.6, tetm=True, potential=True, g_c=0.0, g_pm=0.0, g_r=0.0,
                                             R=0.0, delta_LT=0.0), a, b, V0))
    c0 = T0 @ np.array([ap, am])
    assert abs(c0[0] - cp) > 1e-3 * scale


def test_tetm_planted_slips_are_caught():
    """The closed form with the sign of the (i/2) term flipped, or the kinetic sign flipped, is far
    from what the oracle computes: a slip of either kind in the oracle would fail the test above."""
    p, wave, ap, am, V, V0, a, b, dt = plane_wave_case(3, 2)
    n0 = np.zeros(wave.shape)
    op, _, om, _ = O.rk4_run(p, ap * wave, n0, pot=V, pot_minus=V, dt=dt, steps=1, psi_minus=am * wave, n_minus=n0)
    scale = abs(ap) + abs(am)
    for kw in (dict(i_half_sign=-1.0), dict(kin_sign=-1.0)):
        cp, cm = taylor4(dt * tetm_symbol(p, a, b, V0, **kw)) @ np.array([ap, am])
        assert max(np.abs(op - cp * wave).max(), np.abs(om - cm * wave).max()) > 1e-4 * scale


# ----------------------------------------------------------------------------- TE-TM, zero boundary

def _shift_matrix(W, H, dr, dc):
    """S @ f gives f[r+dr, c+dc] on a W x H row-major lattice, 0 outside."""
    n = W * H
    S = np.zeros((n, n))
    for r in range(H):
        for c in range(W):
            rr, cc = r + dr, c + dc
            if 0 <= rr < H and 0 <= cc < W:
                S[r * W + c, rr * W + cc] = 1.0
    return S


def test_tetm_zero_boundary_equals_padded_lattice_taylor_polynomial():
    Nc, Nr, pad, dt = 7, 6, 4, 0.04
    p = O.params(Nc, Nr, 3.0, 3.0 * 1.3, periodic=(False, False), tetm=True, potential=True, g_c=0.0, g_pm=0.0, g_r=0.0,
                 R=0.0, delta_LT=3e-2)
    rng = np.random.default_rng(11)
    pp = rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc))
    pm = rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc))
    Vp, Vm = 2e-4 * rng.random((Nr, Nc)), 3e-4 * rng.random((Nr, Nc))
    n0 = np.zeros((Nr, Nc))
    op, _, om, _ = O.rk4_run(p, pp, n0, pot=Vp, pot_minus=Vm, dt=dt, steps=1, psi_minus=pm, n_minus=n0)
    W, H = Nc + 2 * pad, Nr + 2 * pad
    E, Wst = _shift_matrix(W, H, 0, 1), _shift_matrix(W, H, 0, -1)
    Nn, Sn = _shift_matrix(W, H, 1, 0), _shift_matrix(W, H, -1, 0)
    hor = (E + Wst) / p.dx ** 2
    ver = (Nn + Sn) / p.dy ** 2
    diag = _shift_matrix(W, H, 1, 1) + _shift_matrix(W, H, -1, -1) - _shift_matrix(W, H, 1, -1) - _shift_matrix(W, H, -1, 1)
    I = np.eye(W * H)
    lap = -2 * (1 / p.dx ** 2 + 1 / p.dy ** 2) * I + hor + ver
    hb = p.h_bar_s
    mix = 0.5j / (p.dx * p.dy) * diag
    Dp = -1j / hb * (p.m_eff_scaled * lap + np.diag(np.pad(Vp, pad).ravel())) - p.gamma_c / 2 * I
    Dm = -1j / hb * (p.m_eff_scaled * lap + np.diag(np.pad(Vm, pad).ravel())) - p.gamma_c / 2 * I
    Op = -1j / hb * p.delta_LT * (hor - ver + mix)
    Om = -1j / hb * p.delta_LT * (hor - ver - mix)
    A = np.block([[Dp, Op], [Om, Dm]])
    y = taylor4(dt * A) @ np.concatenate([np.pad(pp, pad).ravel(), np.pad(pm, pad).ravel()])
    rp = y[:W * H].reshape(H, W)[pad:-pad, pad:-pad]
    rm = y[W * H:].reshape(H, W)[pad:-pad, pad:-pad]
    s = np.abs(rp).max()
    assert np.abs(op - rp).max() < 1e-12 * s
    assert np.abs(om - rm).max() < 1e-12 * s


# ----------------------------------------------------------------------------- TE-TM, uniform spinor

def test_tetm_uniform_spinor_follows_ode():
    """g_pm, reservoirs, an elliptic pump (P+ != P-) and potentials on a uniform spinor: the textbook RK4
    map of (psi+, psi-, n+, n-), with the cross term (2/dx^2 - 2/dy^2) psi-+ of a uniform field."""
    N, dt, steps = 8, 0.04, 5
    p = O.params(N, N, 3.5, 4.9, tetm=True, pump=True, potential=True, g_pm=-2e-5, g_c=6e-5, g_r=1e-4,
                 delta_LT=2.5e-5)
    hb = p.h_bar_s
    Pth = p.gamma_c * p.gamma_r / p.R
    chi = np.pi / 8
    Pp, Pm = 1.5 * Pth * (1 + np.sin(2 * chi)) / 2, 1.5 * Pth * (1 - np.sin(2 * chi)) / 2
    Vp, Vm = 2e-4, -1e-4
    c0 = 2 / p.dx ** 2 - 2 / p.dy ** 2
    assert abs(c0) > 1.0

    def f(y):
        sp, sm, npl, nmi = y
        dsp = -1j / hb * (p.g_c * abs(sp) ** 2 + p.g_r * npl + p.g_pm * abs(sm) ** 2 + Vp) * sp \
            + 0.5 * (p.R * npl - p.gamma_c) * sp - 1j / hb * p.delta_LT * c0 * sm
        dsm = -1j / hb * (p.g_c * abs(sm) ** 2 + p.g_r * nmi + p.g_pm * abs(sp) ** 2 + Vm) * sm \
            + 0.5 * (p.R * nmi - p.gamma_c) * sm - 1j / hb * p.delta_LT * c0 * sp
        dnp = -(p.gamma_r + p.R * abs(sp) ** 2) * npl + Pp
        dnm = -(p.gamma_r + p.R * abs(sm) ** 2) * nmi + Pm
        return np.array([dsp, dsm, dnp, dnm])

    y = np.array([3.0 + 1.0j, -1.0 + 2.0j, 10.0, 7.0], dtype=complex)
    full = lambda v: np.full((N, N), v)
    out = O.rk4_run(p, full(y[0]), full(y[2].real), pump=full(Pp), pot=full(Vp), dt=dt, steps=steps,
                    psi_minus=full(y[1]), n_minus=full(y[3].real), pump_minus=full(Pm), pot_minus=full(Vm))
    for _ in range(steps):
        y = rk4_map(f, y, dt)
    op, onp, om, onm = out
    for got, want in ((op, y[0]), (om, y[1]), (onp, y[2].real), (onm, y[3].real)):
        assert np.abs(got - want).max() < 1e-12 * max(1.0, abs(want)), (got.flat[0], want)
    # each physical coupling moves the result well beyond the tolerance
    assert abs(y[1] - 2.0j + 1.0) > 1e-3


# ----------------------------------------------------------------------------- stochastic terms

def _xi(p, dt, step, N):
    a = p.stochastic_amplitude * np.sqrt(dt)
    return np.array([O.noise_normal2(p.noise_seed, step, c) for c in range(N)]) * a


def _stochastic_rk4(p, psi0, n0, P, xi, dt, tetm_minus=None):
    """Per-cell textbook RK4 with the reference's additive stage noise and drift terms."""
    hb, dV = p.h_bar_s, p.dV
    s = np.sqrt((p.R * n0 + p.gamma_c) / (4 * dV))
    inc = xi * s

    def f(y):
        psi, n = y
        kp = -1j / hb * (p.g_c * abs(psi) ** 2 + p.g_r * n) * psi + 0.5 * (p.R * n - p.gamma_c) * psi \
            - p.g_c * (-1j * psi) / (hb * dV)
        kn = -(p.gamma_r + p.R * abs(psi) ** 2) * n + P + p.R * n / dV
        return np.array([kp, kn])

    y = np.array([psi0, n0 + 0j])
    k1 = f(y)
    y2 = y + dt / 2 * k1
    y2[0] += 0.5 * inc
    k2 = f(y2)
    y3 = y + dt / 2 * k2
    y3[0] += 0.5 * inc
    k3 = f(y3)
    y4 = y + dt * k3
    y4[0] += 1.0 * inc
    k4 = f(y4)
    out = y + dt / 6 * k1 + dt / 3 * k2 + dt / 3 * k3 + dt / 6 * k4
    out[0] += 1.0 * inc
    return out[0], out[1].real


def test_rk_noise_increment_and_drift_terms():
    N, dt, step = 6, 0.04, 17
    # m_eff = 1e30 makes m~ ~ 1e-37: every cell evolves as its own ODE
    p = O.params(N, N, 2.5, 2.5, pump=True, stochastic=0.7, noise_seed=4242, m_eff=1e30)
    assert abs(p.m_eff_scaled) < 1e-30
    rng = np.random.default_rng(5)
    psi0 = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    n0 = 2 + rng.random((N, N))
    P = 1 + rng.random((N, N))
    out, nout = O.rk4_run(p, psi0, n0, pump=P, dt=dt, steps=1, step0=step)
    xi = _xi(p, dt, step, N * N).reshape(N, N)
    want_psi, want_n = _stochastic_rk4(p, psi0, n0, P, xi, dt)
    assert np.abs(out - want_psi).max() < 1e-12 * np.abs(want_psi).max()
    assert np.abs(nout - want_n).max() < 1e-12 * np.abs(want_n).max()
    # the noise is a real part of the result: the same map without it is far away
    dry_psi, _ = _stochastic_rk4(p, psi0, n0, P, 0 * xi, dt)
    assert np.abs(out - dry_psi).max() > 1e-2
    # planted slips: the step-end reservoir in the amplitude, or the drift sign flipped
    s_end = np.sqrt((p.R * want_n + p.gamma_c) / (4 * p.dV))
    s0 = np.sqrt((p.R * n0 + p.gamma_c) / (4 * p.dV))
    slip_psi = want_psi + xi * (s_end - s0)
    assert np.abs(out - slip_psi).max() > 1e-6


def test_rk_noise_other_step_draws_other_numbers():
    N, dt = 4, 0.04
    p = O.params(N, N, 1.5, 1.5, pump=True, stochastic=1.0, noise_seed=9, m_eff=1e30)
    psi0 = np.full((N, N), 1.0 + 0j)
    n0 = np.full((N, N), 2.0)
    P = np.full((N, N), 1.0)
    a, _ = O.rk4_run(p, psi0, n0, pump=P, dt=dt, steps=1, step0=3)
    b, _ = O.rk4_run(p, psi0, n0, pump=P, dt=dt, steps=1, step0=4)
    for st, got in ((3, a), (4, b)):
        want, _ = _stochastic_rk4(p, psi0, n0, P, _xi(p, dt, st, N * N).reshape(N, N), dt)
        assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
    assert np.abs(a - b).max() > 1e-2


def test_tetm_noise_uses_plus_reservoir_for_both_components():
    """With -tetm both components receive the same xi scaled by the psi+ reservoir of the step start."""
    N, dt, step = 4, 0.04, 2
    p = O.params(N, N, 1.5, 1.5, tetm=True, pump=True, stochastic=0.5, noise_seed=77, m_eff=1e30, g_pm=0.0,
                 delta_LT=0.0)
    rng = np.random.default_rng(8)
    pp = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    pm = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    npl, nmi = 1 + rng.random((N, N)), 4 + rng.random((N, N))
    Pp, Pm = np.full((N, N), 1.0), np.full((N, N), 2.0)
    op, onp, om, onm = O.rk4_run(p, pp, npl, pump=Pp, pump_minus=Pm, dt=dt, steps=1, step0=step, psi_minus=pm,
                                 n_minus=nmi)
    xi = _xi(p, dt, step, N * N).reshape(N, N)
    # each component is the scalar map with its own reservoir in the RHS but n+ in the noise amplitude
    s_plus = np.sqrt((p.R * npl + p.gamma_c) / (4 * p.dV))
    xi_m = xi * s_plus / np.sqrt((p.R * nmi + p.gamma_c) / (4 * p.dV))
    wp, wnp = _stochastic_rk4(p, pp, npl, Pp, xi, dt)
    wm, wnm = _stochastic_rk4(p, pm, nmi, Pm, xi_m, dt)
    assert np.abs(op - wp).max() < 1e-12 * np.abs(wp).max()
    assert np.abs(om - wm).max() < 1e-12 * np.abs(wm).max()
    assert np.abs(onp - wnp).max() < 1e-12 * np.abs(wnp).max()
    assert np.abs(onm - wnm).max() < 1e-12 * np.abs(wnm).max()
