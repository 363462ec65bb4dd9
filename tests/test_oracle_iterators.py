"""Pins of the oracle's generic iterator driver (orc_iter_*), CPU only.

The driver restates the reference iterators' step() sequences as data.  It is
checked against what does not come from that transcription:
  - each textbook tableau below satisfies the Butcher order conditions of its
    published order (rooted trees up to order 5, elementary weights computed from
    A and b) -- this pins the coefficients independently of how they were typed;
  - the reference's own coefficient sets are the textbook ones with a short list of
    documented differences (BoSha a32, CashKarp a41 and its K5 error weight,
    Fehlberg5's K5 error weight, the envelope weights c of SSPRK3 and Ralston4);
  - on a uniform periodic field the grid step reduces to the three-variable ODE of
    the condensate/reservoir model, and one oracle step must equal the textbook
    Runge-Kutta map of that tableau (stage equations in matrix form, numpy) with
    time-dependent pump and potential amplitudes, for the fixed and the adaptive
    forms, including the embedded error sum and the norm;
  - the driver reproduces orc_rk4_run and orc_dp45_attempt bitwise;
  - results do not depend on the subgrid decomposition or on the halo width; with
    zero boundaries a linear step equals the stability function of the tableau
    applied to the zero-padded lattice, cropped;
  - a planted transcription slip (one weight changed) fails the uniform-field check.
"""

import numpy as np
import pytest

from oracle import oracle as O

# ----------------------------------------------------------------------------- rooted trees


def _trees_upto(pmax):
    """Rooted trees as sorted tuples of child trees, grouped by order 1..pmax."""
    by_order = {1: [()]}
    allt = [()]

    def multisets(total, min_idx):
        if total == 0:
            yield ()
            return
        for k in range(min_idx, len(allt)):
            t = allt[k]
            o = _order(t)
            if o <= total:
                for rest in multisets(total - o, k):
                    yield (t,) + rest

    for n in range(2, pmax + 1):
        new = sorted(set(tuple(sorted(m, key=repr)) for m in multisets(n - 1, 0)), key=repr)
        by_order[n] = new
        allt.extend(new)
    return by_order


def _order(t):
    return 1 + sum(_order(c) for c in t)


def _gamma(t):
    g = _order(t)
    for c in t:
        g *= _gamma(c)
    return g


def _g(t, A):
    v = np.ones(A.shape[0])
    for c in t:
        v = v * (A @ _g(c, A))
    return v


TREES = _trees_upto(6)


def test_tree_counts():
    # number of rooted trees of order 1..6 (OEIS A000081)
    assert [len(TREES[n]) for n in range(1, 7)] == [1, 1, 2, 4, 9, 20]


def order_defect(A, b, p):
    """max |b.g(t) - 1/gamma(t)| over trees of order <= p."""
    A, b = np.asarray(A, float), np.asarray(b, float)
    return max(abs(b @ _g(t, A) - 1.0 / _gamma(t)) for n in range(1, p + 1) for t in TREES[n])


# ----------------------------------------------------------------------------- textbook tableaux
# (A, b, order of b, b_star or None, order of b_star); sources: Butcher, Numerical Methods for
# Ordinary Differential Equations (2008); Hairer, Norsett, Wanner, Solving ODEs I (1993);
# Ralston, Math. Comp. 16 (1962); Shu & Osher, JCP 77 (1988); Bogacki & Shampine, Appl. Math.
# Lett. 2 (1989); Cash & Karp, ACM TOMS 16 (1990); Fehlberg, NASA TR R-315 (1969);
# Dormand & Prince, J. Comput. Appl. Math. 6 (1980).
def _lower(rows, s):
    A = np.zeros((s, s))
    for i, r in enumerate(rows, start=1):
        A[i, :len(r)] = r
    return A


TEXTBOOK = {
    "Newton": (_lower([], 1), [1.0], 1, None, 0),
    "MP": (_lower([[1 / 2]], 2), [0.0, 1.0], 2, None, 0),
    "Heun2": (_lower([[1.0]], 2), [1 / 2, 1 / 2], 2, None, 0),
    "Ralston": (_lower([[2 / 3]], 2), [1 / 4, 3 / 4], 2, None, 0),
    "Heun3": (_lower([[1 / 3], [0, 2 / 3]], 3), [1 / 4, 0, 3 / 4], 3, None, 0),
    "Ralston3": (_lower([[1 / 2], [0, 3 / 4]], 3), [2 / 9, 1 / 3, 4 / 9], 3, None, 0),
    "RK3": (_lower([[1 / 2], [-1, 2]], 3), [1 / 6, 2 / 3, 1 / 6], 3, None, 0),
    "SSPRK3": (_lower([[1], [1 / 4, 1 / 4]], 3), [1 / 6, 1 / 6, 2 / 3], 3, None, 0),
    "HW": (_lower([[8 / 15], [1 / 4, 5 / 12]], 3), [1 / 4, 0, 3 / 4], 3, None, 0),
    "RK4": (_lower([[1 / 2], [0, 1 / 2], [0, 0, 1]], 4), [1 / 6, 1 / 3, 1 / 3, 1 / 6], 4, None, 0),
    "Rule38": (_lower([[1 / 3], [-1 / 3, 1], [1, -1, 1]], 4), [1 / 8, 3 / 8, 3 / 8, 1 / 8], 4, None, 0),
    "Ralston4": (_lower([[0.4], [0.29697761, 0.15875964], [0.21810040, -3.05096516, 3.83286476]], 4),
                 [0.17476028, -0.55148066, 1.20553560, 0.17118478], 4, None, 0),
    "Nystroem": (_lower([[1 / 3], [4 / 25, 6 / 25], [1 / 4, -3, 15 / 4], [2 / 27, 10 / 9, -50 / 81, 8 / 81],
                         [2 / 25, 12 / 25, 2 / 15, 8 / 75]], 6),
                 [23 / 192, 0, 125 / 192, 0, -27 / 64, 125 / 192], 5, None, 0),
    "Fehlberg2": (_lower([[1 / 2], [1 / 256, 255 / 256]], 3), [1 / 512, 255 / 256, 1 / 512], 2,
                  [1 / 256, 255 / 256, 0], 1),
    "BoSha": (_lower([[1 / 2], [0, 3 / 4], [2 / 9, 1 / 3, 4 / 9]], 4), [2 / 9, 1 / 3, 4 / 9, 0], 3,
              [7 / 24, 1 / 4, 1 / 3, 1 / 8], 2),
    "CashKarp": (_lower([[1 / 5], [3 / 40, 9 / 40], [3 / 10, -9 / 10, 6 / 5], [-11 / 54, 5 / 2, -70 / 27, 35 / 27],
                         [1631 / 55296, 175 / 512, 575 / 13824, 44275 / 110592, 253 / 4096]], 6),
                 [37 / 378, 0, 250 / 621, 125 / 594, 0, 512 / 1771], 5,
                 [2825 / 27648, 0, 18575 / 48384, 13525 / 55296, 277 / 14336, 1 / 4], 4),
    "Fehlberg5": (_lower([[1 / 4], [3 / 32, 9 / 32], [1932 / 2197, -7200 / 2197, 7296 / 2197],
                          [439 / 216, -8, 3680 / 513, -845 / 4104], [-8 / 27, 2, -3544 / 2565, 1859 / 4104, -11 / 40]], 6),
                  [16 / 135, 0, 6656 / 12825, 28561 / 56430, -9 / 50, 2 / 55], 5,
                  [25 / 216, 0, 1408 / 2565, 2197 / 4104, -1 / 5, 0], 4),
    "DP45": (_lower([[1 / 5], [3 / 40, 9 / 40], [44 / 45, -56 / 15, 32 / 9],
                     [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729],
                     [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656],
                     [35 / 384, 0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84]], 7),
             [35 / 384, 0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84, 0], 5,
             [5179 / 57600, 0, 7571 / 16695, 393 / 640, -92097 / 339200, 187 / 2100, 1 / 40], 4),
}


@pytest.mark.parametrize("key", list(TEXTBOOK))
def test_textbook_tableau_has_its_order(key):
    A, b, p, bs, ps = TEXTBOOK[key]
    tol = 5e-8 if key == "Ralston4" else 1e-13  # Ralston's coefficients are published to 8 digits
    assert order_defect(A, b, p) < tol
    assert order_defect(A, b, p + 1) > 1e-4  # and not more
    if bs is not None:
        assert order_defect(A, bs, ps) < tol
        assert order_defect(A, bs, ps + 1) > 1e-5


def test_ralston4_reference_decimals_keep_order_four():
    # the reference passes the same tableau to 11 digits (ralston.cu:111-123)
    A = _lower([[0.4], [0.29697760924, 0.15875964497], [0.21810038822, -3.05096514869, 3.83286476047]], 4)
    b = [0.17476028226, -0.55148066287, 1.2055355994, 0.17118478122]
    assert order_defect(A, b, 4) < 2e-10


def reference_tableau(key):
    """The coefficient set the reference's step() uses: the textbook one with its documented
    differences.  Returns (A, b, c, e) with e None for fixed-step-only iterators; A, c cover
    the stages the adaptive form evaluates (b padded with zeros)."""
    A, b, _, bs, _ = TEXTBOOK[key]
    A = A.copy()
    b = np.array(b, float)
    c = A.sum(axis=1)
    if key == "Ralston4":  # the reference's decimals (ralston.cu:111-123) and its stage weights
        A = _lower([[0.4], [0.29697760924, 0.15875964497], [0.21810038822, -3.05096514869, 3.83286476047]], 4)
        b = np.array([0.17476028226, -0.55148066287, 1.2055355994, 0.17118478122])
        c = np.array([0.0, 0.4, 0.51823725421, 0.51823725421])
    if key == "SSPRK3":
        c[2] = 1.0  # SSPRK3.cu:38
    if key == "BoSha":
        A[2, 1] = 3 / 9  # bogacki.cu:41
        A[3, :3] = b[:3]
        c = A.sum(axis=1)
        c[2] = 3 / 4  # the stage weight passed stays 3/4 (bogacki.cu:43)
    if key == "CashKarp":
        A[3, 0] = 3.0  # cash_karp.cu:47
        c = np.array([0, 1 / 5, 3 / 10, 3 / 5, 1, 7 / 8])
        bs = list(bs)
        bs[4] = 2187 / 6784  # cash_karp.cu:70
    if key == "Fehlberg5":
        bs = list(bs)
        bs[4] = 1 / 5  # fehlberg.cu:137 subtracts +1/5
    e = None if bs is None else b - np.asarray(bs, float)
    return A, b, c, e


def test_reference_deviations_change_the_order():
    # the documented differences are not harmless: they lower the order of the solution
    # (BoSha with a32 = 3/9 keeps only first order: b.c = 7/27, not 1/2)
    A, b, _, _ = reference_tableau("BoSha")
    assert order_defect(A[:3, :3], b[:3], 1) < 1e-13 and order_defect(A[:3, :3], b[:3], 2) > 1e-1
    A, b, _, _ = reference_tableau("CashKarp")
    assert order_defect(A, b, 3) > 1e-2


# ----------------------------------------------------------------------------- uniform field

def uniform_rhs(p, P, V, amps):
    """Condensate/reservoir model on a uniform field (no Laplacian), state (psi, n):
    psi' = -i/hbar (g_c|psi|^2 + g_r n + V a_V) psi + (R n - gamma_c)/2 psi,
    n'   = -(gamma_r + R|psi|^2) n + P a_P,
    with a(c) = a(t) (1 - c) + c a(t + dt)."""
    (pa, pn), (va, vn) = amps

    def F(y, c):
        psi, n = y
        aP = pa * (1 - c) + c * pn
        aV = va * (1 - c) + c * vn
        rho = abs(psi) ** 2
        dpsi = -1j / p.h_bar_s * (p.g_c * rho + p.g_r * n + V * aV) * psi + 0.5 * (p.R * n - p.gamma_c) * psi
        dn = -(p.gamma_r + p.R * rho) * n + P * aP
        return np.array([dpsi, dn])

    return F


def rk_map(A, b, c, y0, h, F):
    """Textbook explicit Runge-Kutta step: Y_i = y0 + h sum_j a_ij K_j, K_i = F(Y_i, c_i)."""
    s = A.shape[0]
    K = []
    for i in range(s):
        Y = y0 + h * sum(A[i, j] * K[j] for j in range(i))
        K.append(F(Y, c[i]))
    out = y0 + h * sum(b[j] * K[j] for j in range(len(b)))
    return out, K


AMPS = ((1.3, 0.7), (0.9, 1.6))


def uniform_case(seed=0):
    N, L = 8, 3.5
    p = O.params(N, N, L, L, pump=True, potential=True, pump_amp=AMPS[0][0], pump_amp_next=AMPS[0][1],
                 pot_amp=AMPS[1][0], pot_amp_next=AMPS[1][1])
    rng = np.random.default_rng(seed)
    psi0 = complex(rng.standard_normal(), rng.standard_normal())
    n0, P, V = rng.random() * 3, 40 * rng.random(), 0.01 * (1 + rng.random())  # |z| = O(1) at dt 0.05
    F = uniform_rhs(p, P, V, AMPS)
    full = lambda v: np.full((N, N), v)
    return p, psi0, n0, P, V, F, full, N


@pytest.mark.parametrize("key", O.ITERATORS)
def test_uniform_field_fixed_step_is_the_textbook_map(key):
    p, psi0, n0, P, V, F, full, N = uniform_case(1)
    A, b, c, e = reference_tableau(key)
    h = 0.05
    # the fixed-step form evaluates only the stages b uses
    s = len(b) if key != "BoSha" else 3
    ref, _ = rk_map(A[:s, :s], b[:s], c[:s], np.array([psi0, n0]), h, F)
    out, n = O.iter_run(p, key, full(psi0), full(n0), pump=full(P), pot=full(V), dt=h, steps=1)
    assert np.abs(out - ref[0]).max() < 1e-13 * abs(ref[0])
    assert np.abs(n - ref[1].real).max() < 1e-13 * abs(ref[1])


@pytest.mark.parametrize("key", [k for k in O.ITERATORS if O.iter_info(k)[1] > 0])
def test_uniform_field_adaptive_attempt_is_the_textbook_map(key):
    p, psi0, n0, P, V, F, full, N = uniform_case(2)
    A, b, c, e = reference_tableau(key)
    h = 0.05
    ref, K = rk_map(A, b, c, np.array([psi0, n0]), h, F)
    out, n, err, norm = O.iter_attempt(p, key, full(psi0), full(n0), pump=full(P), pot=full(V), dt=h)
    assert np.abs(out - ref[0]).max() < 1e-13 * abs(ref[0])
    assert np.abs(n - ref[1].real).max() < 1e-13 * abs(ref[1])
    e_psi = h * sum(e[j] * K[j][0] for j in range(len(e)))
    assert abs(err - N * N * abs(e_psi) ** 2) < 1e-10 * N * N * abs(e_psi) ** 2
    assert abs(norm - N * N * abs(ref[0]) ** 2) < 1e-13 * N * N * abs(ref[0]) ** 2


def test_planted_weight_slip_is_caught():
    """Changing one weight of the reference tableau (a plausible transcription slip) breaks the
    uniform-field identity by far more than its tolerance."""
    p, psi0, n0, P, V, F, full, N = uniform_case(1)
    A, b, c, _ = reference_tableau("Nystroem")
    A2 = A.copy()
    A2[4, 2] = 50 / 81  # sign slip
    ref2, _ = rk_map(A2, b, c, np.array([psi0, n0]), 0.05, F)
    out, _ = O.iter_run(p, "Nystroem", full(psi0), full(n0), pump=full(P), pot=full(V), dt=0.05, steps=1)
    assert np.abs(out - ref2[0]).max() > 1e-6 * abs(ref2[0])


# ----------------------------------------------------------------------------- grid properties

def grid_case(N=24, L=11.5, seed=4):
    rng = np.random.default_rng(seed)
    psi = 0.3 * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
    n, P, V = rng.random((N, N)), 20 * rng.random((N, N)), 1e-3 * rng.random((N, N))
    return psi, n, P, V


def test_generic_driver_reproduces_rk4_and_dp45():
    psi, n, P, V = grid_case()
    for periodic in (True, False):
        p = O.params(24, 24, 11.5, 11.5, periodic=(periodic, periodic), pump=True, potential=True)
        a = O.rk4_run(p, psi, n, pump=P, pot=V, dt=0.03, steps=4)
        b = O.iter_run(p, "RK4", psi, n, pump=P, pot=V, dt=0.03, steps=4)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        c = O.dp45_attempt(p, psi, n, pump=P, pot=V, dt=0.03, step=3)
        d = O.iter_attempt(p, "DP45", psi, n, pump=P, pot=V, dt=0.03, step=3)
        assert np.array_equal(c[0], d[0]) and np.array_equal(c[1], d[1]) and c[2] == d[2] and c[3] == d[3]
    # stochastic and TE-TM paths as well
    ps = O.params(24, 24, 11.5, 11.5, pump=True, stochastic=0.05, noise_seed=9)
    a = O.rk4_run(ps, psi, n, pump=P, dt=0.03, steps=3, step0=5)
    b = O.iter_run(ps, "RK4", psi, n, pump=P, dt=0.03, steps=3, step0=5)
    assert np.array_equal(a[0], b[0])
    pt = O.params(24, 24, 11.5, 11.5, pump=True, tetm=True)
    a = O.rk4_run(pt, psi, n, pump=P, dt=0.03, steps=2, psi_minus=psi[::-1], n_minus=n.T, pump_minus=P.T)
    b = O.iter_run(pt, "RK4", psi, n, pump=P, dt=0.03, steps=2, psi_minus=psi[::-1], n_minus=n.T, pump_minus=P.T)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("key", ["MP", "SSPRK3", "Rule38", "Nystroem", "CashKarp"])
@pytest.mark.parametrize("periodic", [True, False])
def test_subgrid_decomposition_invariance(key, periodic):
    psi, n, P, V = grid_case()
    outs = []
    for sg in [(1, 1), (2, 2), (4, 3)]:
        p = O.params(24, 24, 11.5, 11.5, periodic=(periodic, periodic), subgrids=sg, pump=True, potential=True)
        outs.append(O.iter_run(p, key, psi, n, pump=P, pot=V, dt=0.03, steps=3))
    for o in outs[1:]:
        assert np.array_equal(o[0], outs[0][0]) and np.array_equal(o[1], outs[0][1])


def stability_matrix_function(A, b, M):
    """R(M) = I + sum_k (b^T A^k 1) M^(k+1) for an explicit tableau (A nilpotent)."""
    s = A.shape[0]
    out = np.eye(M.shape[0], dtype=complex)
    Mk = np.eye(M.shape[0], dtype=complex)
    Ak1 = np.ones(s)
    for _ in range(s):
        Mk = Mk @ M
        out = out + (b @ Ak1) * Mk
        Ak1 = A @ Ak1
    return out


def laplacian(W, H, dx, dy):
    n = W * H
    Lp = np.zeros((n, n))
    for r in range(H):
        for c in range(W):
            i = r * W + c
            Lp[i, i] = -2 / dx ** 2 - 2 / dy ** 2
            for dr, dc, w in ((0, 1, 1 / dx ** 2), (0, -1, 1 / dx ** 2), (1, 0, 1 / dy ** 2), (-1, 0, 1 / dy ** 2)):
                rr, cc = r + dr, c + dc
                if 0 <= rr < H and 0 <= cc < W:
                    Lp[i, rr * W + cc] += w
    return Lp


@pytest.mark.parametrize("key", ["Heun2", "RK3", "Ralston4", "Nystroem", "Fehlberg5"])
def test_zero_boundary_linear_step_is_padded_stability_function(key):
    """Zero boundaries: the halo ring evolves within the step, so one step equals the
    tableau's stability function on the lattice zero-padded by the number of stages, cropped
    (the halo width of the reference may be larger: the result must not depend on it)."""
    Nc, Nr, L, dt = 9, 7, 4.0, 0.05
    p = O.params(Nc, Nr, L, L * (Nr - 1) / (Nc - 1), periodic=(False, False), reservoir=False, potential=True, g_c=0.0)
    rng = np.random.default_rng(8)
    psi = rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc))
    V = 2e-4 * rng.random((Nr, Nc))
    out, _ = O.iter_run(p, key, psi, pot=V, dt=dt, steps=1)
    A, b, _, _ = reference_tableau(key)
    pad = len(b)
    W, H = Nc + 2 * pad, Nr + 2 * pad
    Vp = np.pad(V, pad)
    M = -1j / p.h_bar_s * (p.m_eff_scaled * laplacian(W, H, p.dx, p.dy) + np.diag(Vp.ravel())) - p.gamma_c / 2 * np.eye(W * H)
    ref = (stability_matrix_function(A[:len(b), :len(b)], b, dt * M) @ np.pad(psi, pad).ravel()).reshape(H, W)
    ref = ref[pad:-pad, pad:-pad]
    assert np.abs(out - ref).max() < 1e-12 * np.abs(ref).max()


def test_adaptive_controller_power():
    # BoSha adapts with exponent 1/3 (bogacki.cu:69), the others with 1/4
    acc, nd, fe = O.adapt_dt_pow(1e-8, 1.0, 1e-6, 0.01, 1e-4, 0.1, 1 / 3)
    assert acc and abs(nd - 0.01 * (1e-6 / 2 / 1e-8) ** (1 / 3)) < 1e-15
    assert O.iter_info("BoSha")[1] == pytest.approx(1 / 3)
    assert all(O.iter_info(k)[1] == 0.25 for k in ("DP45", "CashKarp", "Fehlberg2", "Fehlberg5"))


@pytest.mark.parametrize("key", ["BoSha", "CashKarp", "Fehlberg2", "Fehlberg5"])
def test_adaptive_run_reaches_tmax_and_is_consistent(key):
    """An adaptive run is a sequence of accepted attempts: replaying the attempts with the
    returned dt sequence reproduces it, and every accepted error is below the tolerance."""
    N, L = 16, 7.5
    p = O.params(N, N, L, L, pump=True)
    psi, n, P, _ = grid_case(N, L, 11)
    res = O.iter_adaptive_run(p, key, psi, n, pump=P, t=0.0, dt=0.02, t_max=0.6, tol=1e-7, dt_min=1e-5, dt_max=0.1)
    assert res["t"] >= 0.6 and res["accepted"] > 3
    # replay
    t, dt, x, y = 0.0, 0.02, psi.copy(), n.copy()
    steps = 0
    while t < 0.6:
        t += dt
        while True:
            a, b_, err, nrm = O.iter_attempt(p, key, x, y, pump=P, dt=dt, step=steps)
            acc, dt, fe = O.adapt_dt_pow(err, nrm, 1e-7, dt, 1e-5, 0.1, O.iter_info(key)[1])
            if acc:
                assert fe < 1e-7
                break
        x, y = a, b_
        steps += 1
    assert steps == res["accepted"]
    assert np.array_equal(x, res["psi"]) and np.array_equal(y, res["n"])
