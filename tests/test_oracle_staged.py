"""Pins of the oracle's 12- and 13-stage iterators and of its source chains, CPU only.

  - NSRK78 (NSRK78.cu:22-116) and DP85 (dormand_prince.cu:128-205): the coefficient sets
    the reference passes satisfy the Butcher order conditions (rooted trees up to order 6;
    both outputs are of order 8, NSRK78's embedded solution of order 7, DP85's 5th-order
    weights b5 of order exactly 5), and their envelope weights c are the row sums of A to
    the 10 decimals written there;
  - on a uniform periodic field one oracle step of NSRK78 / DP85 equals the textbook
    Runge-Kutta map of those coefficients on the condensate/reservoir ODE (fixed and
    adaptive forms, error sum and norm included);
  - the source chains: pump, potential and pulse as sums over temporal groups with complex
    temporal amplitudes (kernel_gp_compute.cuh:68-100) -- on a uniform field one step equals
    the textbook map of the ODE
        psi' = -i/hbar (g_c|psi|^2 + g_r n + V(c)) psi + (R n - gamma_c)/2 psi + F(c)/hbar,
        n'   = -(gamma_r + R|psi|^2) n + Re P(c),
    X(c) = sum_g X_g (a_g(t) (1 - c) + c a_g(t + dt)), for RK4, DP45 and NSRK78; a single
    group at amplitude 1 reproduces the plain pump / potential bitwise; with no nonlinearity
    the step is affine in the pulse field; a planted slip (conjugated pulse amplitude) fails;
  - subgrid invariance with chains, zero and periodic boundaries.
"""

import numpy as np
import pytest

from oracle import oracle as O
from test_oracle_iterators import order_defect, rk_map

NSRK78_C = [0.0, 0.001, 0.1111111111, 0.1666666667, 0.4166666667, 0.5, 0.8333333333, 0.1666666667, 0.6666666667, 0.76, 0.84, 1.0, 1.0]
NSRK78_A = [
    [1 / 1000.0],
    [-491 / 81.0, 500 / 81.0],
    [1 / 24.0, 0.0, 1 / 8.0],
    [5 / 12.0, 0.0, -25 / 16.0, 25 / 16.0],
    [1 / 20.0, 0.0, 0.0, 1 / 4.0, 1 / 5.0],
    [-43 / 180.0, 0.0, 0.0, 127 / 108.0, -329 / 135.0, 7 / 3.0],
    [27931 / 240300.0, 0.0, 0.0, -631 / 16020.0, 2459 / 6675.0, -3572 / 12015.0, 5 / 267.0],
    [26114 / 12015.0, 0.0, 0.0, -7480 / 801.0, 67264 / 4005.0, -30640 / 2403.0, 1051 / 1335.0, 3.0],
    [33096587331 / 17382812500.0, 0.0, 0.0, -1828977848 / 173828125.0, 62801809904 / 4345703125.0, -9389764774 / 869140625.0, 6380757669 / 8691406250.0, 98417891 / 19531250.0, -1692691 / 39062500.0],
    [-1456295425347 / 2642187500000.0, 0.0, 0.0, 110740056 / 34765625.0, -21221682384 / 4345703125.0, 58859060169 / 13906250000.0, -177381525069 / 1529687500000.0, -28942485159 / 27812500000.0, -1272297 / 312500000.0, 5151 / 297616.0],
    [844300798 / 137013275.0, 0.0, 0.0, -9997568 / 206035.0, 49636624 / 1030175.0, -3358834871 / 91067470.0, -40456983 / 1813108.0, 495817135 / 16647628.0, -149375 / 84266.0, 7470703125 / 1567431866.0, 1562500000 / 72342361.0],
    [-26225423 / 37371100.0, 0.0, 0.0, 807744 / 98345.0, -205584 / 37825.0, 5882202 / 1278485.0, -84543 / 432718.0, -223415 / 39338.0, -3625 / 6188.0, 292968750 / 374084711.0, 0.0, 0.0],
]
NSRK78_B = [4241 / 88200.0, 0.0, 0.0, 0.0, 0.0, 9946 / 23205.0, -10449 / 1925.0, 415449 / 1573075.0, -2025 / 5096.0, 48828125 / 44900856.0, 48828125 / 9843561.0, 463 / 12600.0, 0.0]
NSRK78_E = [4241 / 88200.0 - 3799 / 79800.0, 0.0, 0.0, 0.0, 0.0, 9946 / 23205.0 - 538 / 1365.0, -10449 / 1925.0 - 351 / 1925.0, 415449 / 1573075.0 - 4149 / 15575.0, -2025 / 5096.0 - (-45 / 392), 48828125 / 44900856.0 - 48828125 / 284372088.0, 48828125 / 9843561.0 - 0.0, 463 / 12600.0 - 0.0, 0.0 - 221 / 4200.0]
DP85_C = [0.0, 0.05260015196, 0.07890022794, 0.1183503419, 0.2816496581, 0.3333333333, 0.25, 0.3076923077, 0.6512820513, 0.6, 0.8571428571, 1.0]
DP85_A = [
    [5.260015195876773e-2],
    [1.9725056984537899e-2, 5.9175170953613698e-2],
    [2.9587585476806849e-2, 0.0, 8.8762756430420548e-2],
    [2.4136513415926669e-1, 0.0, -8.8454947932828610e-1, 9.2483400326179200e-1],
    [3.7037037037037037e-2, 0.0, 0.0, 1.7082860872947387e-1, 1.2546768756682243e-1],
    [3.7109375e-2, 0.0, 0.0, 1.7025221101954404e-1, 6.0216538980455961e-2, -1.7578125e-2],
    [3.7092000118504793e-2, 0.0, 0.0, 1.7038392571223999e-1, 1.0726203044637328e-1, -1.5319437748624402e-2, 8.2737891638140229e-3],
    [6.2411095871607572e-1, 0.0, 0.0, -3.3608926294469413e+0, -8.6821934684172601e-1, 2.7592099699446708e+1, 2.0154067550477893e+1, -4.3489884181069959e+1],
    [4.7766253643826437e-1, 0.0, 0.0, -2.4881146199716676e+0, -5.9029082683684300e-1, 2.1230051448181194e+1, 1.5279233632882424e+1, -3.3288210968984863e+1, -2.0331201708508626e-2],
    [-9.3714243008598733e-1, 0.0, 0.0, 5.1863724288440637e+0, 1.0914373489967296e+0, -8.1497870107469261e+0, -1.8520065659996960e+1, 2.2739487099350504e+1, 2.4936055526796524e+0, -3.0467644718982195e+0],
    [2.2733101475165382e+0, 0.0, 0.0, -1.0534495466737250e+1, -2.0008720582248625e+0, -1.7958931863118799e+1, 2.7948884529419960e+1, -2.8589982771350237e+0, -8.8728569335306295e+0, 1.2360567175794303e+1, 6.4339274601576353e-1],
]
DP85_B = [5.4293734116568762e-2, 0.0, 0.0, 0.0, 0.0, 4.4503128927524089e+0, 1.8915178993145004e+0, -5.8012039600105848e+0, 3.1116436695781989e-1, -1.5216094966251608e-1, 2.0136540080403035e-1, 4.4710615727772591e-2]
DP85_B5 = [4.1173689122373888e-2, 0.0, 0.0, 0.0, 0.0, 5.6754693391286133e+0, 2.3872768489717506e+0, -7.4655811424655713e+0, 6.6149321570779360e-1, -4.8634006837553356e-1, 1.1944219431891464e-1, 6.7065923591658886e-2]


def lower(rows, s):
    A = np.zeros((s, s))
    for i, r in enumerate(rows, start=1):
        A[i, :len(r)] = r
    return A


def tableau(key):
    """(A, b, c, e) as the reference passes them; e = b - embedded weights."""
    if key == "NSRK78":
        A, b, c = lower(NSRK78_A, 13), np.array(NSRK78_B), np.array(NSRK78_C)
        return A, b, c, np.array(NSRK78_E)
    A, b, c = lower(DP85_A, 12), np.array(DP85_B), np.array(DP85_C)
    return A, b, c, b - np.array(DP85_B5)


def test_nsrk78_and_dp85_order_conditions():
    for key in ("NSRK78", "DP85"):
        A, b, c, e = tableau(key)
        assert order_defect(A, b, 6) < 1e-11, key
        assert np.abs(A.sum(axis=1) - c).max() < 1e-10, key
    A, b, c, e = tableau("NSRK78")
    assert order_defect(A, b - e, 6) < 1e-11  # 7th-order embedded solution
    A, b, c, e = tableau("DP85")
    assert order_defect(A, b - e, 5) < 1e-11 and order_defect(A, b - e, 6) > 1e-5


def test_oracle_iterator_info():
    assert O.iter_info("NSRK78")[:2] == (13, 1 / 8) and O.iter_info("NSRK78")[2:] == (13, 13)
    assert O.iter_info("DP85")[:2] == (12, 1 / 4) and O.iter_info("DP85")[2:] == (12, 12)


def chain_rhs(p, groups):
    """Uniform-field ODE with source chains; groups[q] = [(value, a0, a1), ...] for q in pump,
    potential, pulse (values complex for the pulse)."""
    def X(q, c):
        return sum(v * (a0 * (1 - c) + c * a1) for v, a0, a1 in groups[q])

    def F(y, c):
        psi, n = y
        rho = abs(psi) ** 2
        V, Fp, P = X("potential", c), X("pulse", c), X("pump", c)
        dpsi = (-1j / p.h_bar_s * (p.g_c * rho + p.g_r * n + V) * psi + 0.5 * (p.R * n - p.gamma_c) * psi
                + Fp / p.h_bar_s)
        dn = -(p.gamma_r + p.R * rho) * n + np.real(P)
        return np.array([dpsi, dn])
    return F


GROUPS = {
    "pump": [(23.0, 1.3, 0.7), (11.0, 0.4, 0.9)],
    "potential": [(0.012, 0.9 + 0.2j, 1.6 - 0.1j), (0.004, 0.3j, -0.5 + 0.25j)],
    "pulse": [(0.002 + 0.001j, 1.1 - 0.4j, 0.8 + 0.6j), (-0.0015j, 0.5 + 0.5j, 0.2 - 0.9j)],
}


def chain_params(N, L, groups, periodic=(True, True), subgrids=(1, 1), **kw):
    p = O.params(N, N, L, L, periodic=periodic, subgrids=subgrids, pump=True, potential=True, **kw)
    return p


def set_uniform_chains(p, N, groups):
    for q, gl in groups.items():
        O.set_chain(p, q, [(np.full((N, N), v), None, a0, a1) for v, a0, a1 in gl])


def reference_tableau_of(key):
    if key in ("NSRK78", "DP85"):
        return tableau(key)
    from test_oracle_iterators import reference_tableau
    return reference_tableau(key)


@pytest.mark.parametrize("key", ["RK4", "DP45", "NSRK78", "DP85"])
def test_uniform_field_chains_fixed_step_is_the_textbook_map(key):
    N, L, h = 16, 7.5, 0.05
    p = chain_params(N, L, GROUPS)
    set_uniform_chains(p, N, GROUPS)
    A, b, c, e = reference_tableau_of(key)
    s = len(b) if key != "DP45" else 6
    psi0, n0 = 0.8 - 0.5j, 1.7
    ref, _ = rk_map(A[:s, :s], b[:s], c[:s], np.array([psi0, n0]), h, chain_rhs(p, GROUPS))
    out, n = O.iter_run(p, key, np.full((N, N), psi0), np.full((N, N), n0), dt=h, steps=1)
    assert np.abs(out - ref[0]).max() < 1e-13 * abs(ref[0]), key
    assert np.abs(n - ref[1].real).max() < 1e-13 * abs(ref[1]), key


@pytest.mark.parametrize("key", ["DP45", "NSRK78", "DP85"])
def test_uniform_field_chains_adaptive_attempt_is_the_textbook_map(key):
    N, L, h = 16, 7.5, 0.05
    p = chain_params(N, L, GROUPS)
    set_uniform_chains(p, N, GROUPS)
    A, b, c, e = reference_tableau_of(key)
    psi0, n0 = 0.6 + 0.9j, 2.3
    ref, K = rk_map(A, b, c, np.array([psi0, n0]), h, chain_rhs(p, GROUPS))
    out, n, err, norm = O.iter_attempt(p, key, np.full((N, N), psi0), np.full((N, N), n0), dt=h)
    assert np.abs(out - ref[0]).max() < 1e-13 * abs(ref[0]), key
    assert np.abs(n - ref[1].real).max() < 1e-13 * abs(ref[1]), key
    e_psi = h * sum(e[j] * K[j][0] for j in range(len(e)))
    assert abs(err - N * N * abs(e_psi) ** 2) < 1e-8 * N * N * abs(e_psi) ** 2, key
    assert abs(norm - N * N * abs(ref[0]) ** 2) < 1e-13 * N * N * abs(ref[0]) ** 2, key


def grid(N=20, seed=3):
    rng = np.random.default_rng(seed)
    psi = 0.3 * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
    n = rng.random((N, N))
    return rng, psi, n


def test_single_unit_group_reproduces_plain_sources_bitwise():
    N, L = 20, 9.5
    rng, psi, n = grid(N)
    P, V = 20 * rng.random((N, N)), 1e-3 * rng.random((N, N))
    for periodic in (True, False):
        p0 = O.params(N, N, L, L, periodic=(periodic, periodic), pump=True, potential=True)
        a = O.iter_run(p0, "RK4", psi, n, pump=P, pot=V, dt=0.03, steps=3)
        p1 = O.params(N, N, L, L, periodic=(periodic, periodic), pump=True, potential=True)
        O.set_chain(p1, "pump", [(P, None, 1.0, 1.0)])
        O.set_chain(p1, "potential", [(V, None, 1.0, 1.0)])
        b = O.iter_run(p1, "RK4", psi, n, dt=0.03, steps=3)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_step_is_affine_in_the_pulse():
    """Without nonlinearity and reservoir the step is an affine map of the pulse field."""
    N, L = 20, 9.5
    rng, psi, _ = grid(N)
    F1 = 1e-3 * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
    F2 = 1e-3 * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))

    def run(F):
        p = O.params(N, N, L, L, periodic=(False, True), reservoir=False, g_c=0.0)
        O.set_chain(p, "pulse", [(F, None, 0.7 + 0.2j, 0.1 - 0.5j)])
        return O.iter_run(p, "DP85", psi, None, dt=0.02, steps=2)[0]

    base = run(np.zeros((N, N), complex))
    d1, d2, d12 = run(F1) - base, run(F2) - base, run(F1 + F2) - base
    assert np.abs(d12 - (d1 + d2)).max() < 1e-12 * np.abs(d12).max()
    assert np.abs(run(2.0 * F1) - base - 2.0 * d1).max() < 1e-12 * np.abs(d1).max()


def test_planted_conjugated_pulse_amplitude_is_caught():
    N, L, h = 16, 7.5, 0.05
    p = chain_params(N, L, GROUPS)
    set_uniform_chains(p, N, GROUPS)
    bad = dict(GROUPS)
    bad["pulse"] = [(v, np.conj(a0), np.conj(a1)) for v, a0, a1 in GROUPS["pulse"]]
    A, b, c, _ = reference_tableau_of("RK4")
    ref_bad, _ = rk_map(A, b, c, np.array([0.8 - 0.5j, 1.7]), h, chain_rhs(p, bad))
    out, _ = O.iter_run(p, "RK4", np.full((N, N), 0.8 - 0.5j), np.full((N, N), 1.7), dt=h, steps=1)
    assert np.abs(out - ref_bad[0]).max() > 1e-9


@pytest.mark.parametrize("periodic", [True, False])
def test_chains_subgrid_invariance(periodic):
    N, L = 48, 23.5  # subgrid edges >= the 12-cell halo of DP85
    rng, psi, n = grid(N, 5)
    groups = {q: [(np.asarray(20 * rng.random((N, N)) if q == "pump" else (1e-3 * rng.random((N, N)) if q == "potential"
                else 1e-3 * (rng.random((N, N)) + 1j * rng.random((N, N))))), None, a0, a1) for _, a0, a1 in gl]
              for q, gl in GROUPS.items()}
    out = []
    for sg in ((1, 1), (2, 2), (4, 2)):
        p = O.params(N, N, L, L, periodic=(periodic, periodic), subgrids=sg, pump=True, potential=True)
        for q, gl in groups.items():
            O.set_chain(p, q, gl)
        out.append(O.iter_run(p, "DP85", psi, n, dt=0.02, steps=2))
    for o in out[1:]:
        assert np.array_equal(o[0], out[0][0]) and np.array_equal(o[1], out[0][1])
