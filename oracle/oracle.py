"""ctypes wrapper of the FP64 CPU oracle (oracle/libphx_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the parity checker.  The product package
(phoenix_amd) never imports this module.

Arrays use the product layout: full grid, row-major (N_r, N_c); psi as complex128, n as float64
(complex128 when the parameters set complex_reservoir), P / V as float64.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libphx_oracle.so")


MAX_GROUPS = 8  # ORC_MAX_GROUPS


class OrcParams(C.Structure):
    _fields_ = [
        ("N_c", C.c_int), ("N_r", C.c_int),
        ("subgrids_columns", C.c_int), ("subgrids_rows", C.c_int),
        ("periodic_x", C.c_int), ("periodic_y", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dV", C.c_double),
        ("one_over_dx2", C.c_double), ("one_over_dy2", C.c_double), ("m2_over_dx2_p_dy2", C.c_double),
        ("m_eff_scaled", C.c_double), ("one_over_h_bar_s", C.c_double), ("h_bar_s", C.c_double), ("m_eff", C.c_double),
        ("gamma_c", C.c_double), ("gamma_r", C.c_double), ("g_c", C.c_double), ("g_r", C.c_double),
        ("R", C.c_double), ("g_pm", C.c_double), ("delta_LT", C.c_double),
        ("tetm", C.c_int), ("use_reservoir", C.c_int), ("use_pump", C.c_int), ("use_potential", C.c_int),
        ("use_stochastic", C.c_int),
        ("stochastic_amplitude", C.c_double),
        ("pump_amp", C.c_double), ("pump_amp_next", C.c_double), ("pot_amp", C.c_double), ("pot_amp_next", C.c_double),
        ("noise_seed", C.c_uint64),
        ("imag_time_amplitude", C.c_double),
        ("L_x", C.c_double), ("L_y", C.c_double),
        ("src_groups", C.c_int * 3),
        ("src_field", C.c_void_p * 2 * MAX_GROUPS * 3),
        ("src_amp", C.c_double * 2 * MAX_GROUPS * 3), ("src_amp_next", C.c_double * 2 * MAX_GROUPS * 3),
        ("complex_reservoir", C.c_int),
    ]


def build() -> str:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        P = C.POINTER(OrcParams)
        L.orc_defaults.argtypes = [P]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads.restype = C.c_int
        L.orc_derive.argtypes = [P] + [C.c_double] * 7
        L.orc_derive.restype = C.c_double
        L.orc_rk4_run.argtypes = [P, dp, dp, dp, dp, dp, dp, dp, dp, C.c_double, C.c_int, C.c_int64]
        L.orc_dp45_attempt.argtypes = [P, dp, dp, dp, dp, dp, dp, dp, dp, C.c_double, C.c_double, C.c_int64, dp, dp]
        L.orc_adapt_dt.argtypes = [C.c_double] * 6 + [dp, dp]
        L.orc_dp45_run.argtypes = [P, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_int, C.POINTER(C.c_int64)]
        L.orc_dp45_run.restype = C.c_int64
        L.orc_fft_filter.argtypes = [dp, dp, C.c_int, C.c_int]
        L.orc_fft2d.argtypes = [dp, C.c_int, C.c_int, C.c_int]
        L.orc_envelope_add.argtypes = [dp, C.c_int, C.c_int] + [C.c_double] * 10 + [C.c_int] + [C.c_double] * 3 + [C.c_int]
        L.orc_noise_normal2.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_int, dp, dp]
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_gaussian_field.argtypes = [dp, C.c_int, C.c_double, C.c_uint64]
        L.orc_rk4_adjoint.argtypes = [P, dp, dp, dp, dp, C.c_double, C.c_int, dp, dp, dp, dp, dp, dp]
        cs = C.c_char_p
        L.orc_iter_info.argtypes = [cs, C.POINTER(C.c_int), dp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_iter_run.argtypes = [P, cs, dp, dp, dp, dp, dp, dp, dp, dp, C.c_double, C.c_int, C.c_int64]
        L.orc_iter_attempt.argtypes = [P, cs, dp, dp, dp, dp, dp, dp, dp, dp, C.c_double, C.c_double, C.c_int64, dp, dp]
        L.orc_adapt_dt_pow.argtypes = [C.c_double] * 7 + [dp, dp]
        L.orc_imag_normalize.argtypes = [P, dp, dp]
        L.orc_imag_normalize_reservoir.argtypes = [P, dp, dp]
        L.orc_ssfm_run.argtypes = [P, dp, dp, dp, dp, dp, dp, dp, dp, C.c_double, C.c_int, C.c_int64]
        L.orc_iter_adaptive_run.argtypes = [P, cs, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, C.c_double, C.c_double,
                                            C.c_double, C.c_double, C.c_int, C.POINTER(C.c_int64)]
        L.orc_iter_adaptive_run.restype = C.c_int64
        _lib = L
    return _lib


def set_threads(n: int = 0) -> int:
    """Threads of the oracle's OpenMP subgrid loops (n <= 0: query); returns the count in use."""
    return int(lib().orc_set_threads(int(n)))


# reference defaults: system.cpp:20-23 (SI constants)
M_E = 9.10938356e-31
H_BAR = 1.0545718e-34
E_E = 1.60217662e-19


def params(N_c, N_r, L_x, L_y, *, periodic=(True, True), subgrids=(1, 1), tetm=False, reservoir=True,
           pump=False, potential=False, stochastic=0.0, noise_seed=0, **phys) -> OrcParams:
    p = OrcParams()
    lib().orc_defaults(C.byref(p))
    p.N_c, p.N_r = N_c, N_r
    p.subgrids_columns, p.subgrids_rows = subgrids
    p.periodic_x, p.periodic_y = int(periodic[0]), int(periodic[1])
    for k, v in phys.items():
        setattr(p, k, v)
    h_bar_s = phys.get("h_bar_s", -1.0)
    m_eff = phys.get("m_eff", -1.0)
    lib().orc_derive(C.byref(p), L_x, L_y, H_BAR, E_E, M_E, h_bar_s, m_eff)
    p.tetm = int(tetm)
    p.use_reservoir = int(reservoir)
    p.use_pump = int(pump)
    p.use_potential = int(potential)
    p.use_stochastic = int(stochastic > 0)
    p.stochastic_amplitude = stochastic
    p.noise_seed = noise_seed
    return p


SOURCES = {"pump": 0, "potential": 1, "pulse": 2}


def set_chain(p: OrcParams, source: str, groups):
    """Source chain of `source` ("pump", "potential", "pulse"): one entry per temporal group,
    (field_plus, field_minus or None, amp, amp_next) with complex temporal amplitudes at t and
    t + dt; pump and potential fields real (N_r, N_c), pulse fields complex.  The arrays are
    kept alive on `p`."""
    q = SOURCES[source]
    keep = p.__dict__.setdefault("_chain_arrays", {})
    keep[q] = []
    p.src_groups[q] = len(groups)
    for g, (fp, fm, a0, a1) in enumerate(groups):
        for comp, f in enumerate((fp, fm)):
            if f is None:
                p.src_field[q][g][comp] = None
                continue
            arr = _c128(f).view(np.float64) if q == 2 else _f64(f)
            arr = np.ascontiguousarray(arr)
            keep[q].append(arr)
            p.src_field[q][g][comp] = arr.ctypes.data
        p.src_amp[q][g][0], p.src_amp[q][g][1] = complex(a0).real, complex(a0).imag
        p.src_amp_next[q][g][0], p.src_amp_next[q][g][1] = complex(a1).real, complex(a1).imag


def _dp(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128).copy()


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64).copy()


def _res(p, n, shape, keep_none=False):
    """A reservoir array in the layout the parameters ask for: complex128 with complex_reservoir,
    float64 otherwise; None stays None with keep_none, else zeros."""
    if n is None:
        if keep_none:
            return None
        n = np.zeros(shape)
    return _c128(n) if p.complex_reservoir else _f64(n)


def _rdp(a):
    if a is None:
        return None
    return _dp(a.view(np.float64)) if np.iscomplexobj(a) else _dp(a)


def rk4_run(p: OrcParams, psi, n=None, pump=None, pot=None, dt=0.04, steps=1, psi_minus=None, n_minus=None,
            pump_minus=None, pot_minus=None, step0=0):
    psi = _c128(psi)
    pm = _c128(psi_minus) if psi_minus is not None else None
    n = _res(p, n, psi.shape)
    nm = _res(p, n_minus, psi.shape) if (n_minus is not None or psi_minus is not None) else None
    pump = _f64(pump)
    pot = _f64(pot)
    rc = lib().orc_rk4_run(C.byref(p), _dp(psi.view(np.float64)), _dp(pm.view(np.float64)) if pm is not None else None,
                           _rdp(n), _rdp(nm), _dp(pump), _dp(_f64(pump_minus)), _dp(pot), _dp(_f64(pot_minus)),
                           dt, steps, step0)
    if rc:
        raise RuntimeError(f"orc_rk4_run failed: {rc}")
    if psi_minus is not None:
        return psi, n, pm, nm
    return psi, n


def rk4_adjoint(p: OrcParams, psi0, n0, pump, pot, dt, steps, lam_psi_T, lam_n_T=None):
    """Discrete adjoint of `steps` RK4 steps (orc_rk4_adjoint).  lam_psi_T: complex cotangent
    (dJ/dRe + i dJ/dIm) of the final psi, lam_n_T of the final n.  Returns
    (lam_psi_0, lam_n_0, dJ/dP, dJ/dV, psi_T, n_T)."""
    shape = np.asarray(psi0).shape
    psi0 = _c128(psi0)
    n0 = _f64(n0) if n0 is not None else np.zeros(shape)
    lp = _c128(lam_psi_T)
    ln = _f64(lam_n_T) if lam_n_T is not None else np.zeros(shape)
    gP, gV = np.zeros(shape), np.zeros(shape)
    psiT, nT = np.zeros(shape, np.complex128), np.zeros(shape)
    rc = lib().orc_rk4_adjoint(C.byref(p), _dp(psi0.view(np.float64)), _dp(n0), _dp(_f64(pump)), _dp(_f64(pot)), dt, steps,
                               _dp(lp.view(np.float64)), _dp(ln), _dp(gP), _dp(gV), _dp(psiT.view(np.float64)), _dp(nT))
    if rc:
        raise RuntimeError(f"orc_rk4_adjoint failed: {rc}")
    return lp, ln, gP, gV, psiT, nT


def dp45_attempt(p, psi, n, pump=None, pot=None, dt=0.04, noise_dt=None, step=0, psi_minus=None, n_minus=None,
                 pump_minus=None, pot_minus=None):
    psi = _c128(psi)
    n = _res(p, n, psi.shape, keep_none=True)
    pm = _c128(psi_minus) if psi_minus is not None else None
    nm = _res(p, n_minus, psi.shape, keep_none=True)
    es, nr = C.c_double(), C.c_double()
    rc = lib().orc_dp45_attempt(C.byref(p), _dp(psi.view(np.float64)), _dp(pm.view(np.float64)) if pm is not None else None,
                                _rdp(n), _rdp(nm), _dp(_f64(pump)), _dp(_f64(pump_minus)), _dp(_f64(pot)),
                                _dp(_f64(pot_minus)), dt, dt if noise_dt is None else noise_dt, step,
                                C.byref(es), C.byref(nr))
    if rc:
        raise RuntimeError("orc_dp45_attempt failed")
    out = (psi, n, es.value, nr.value)
    if psi_minus is not None:
        out = out + (pm, nm)
    return out


def adapt_dt(err_sum, norm, tol, dt, dt_min, dt_max):
    nd, fe = C.c_double(), C.c_double()
    acc = lib().orc_adapt_dt(err_sum, norm, tol, dt, dt_min, dt_max, C.byref(nd), C.byref(fe))
    return bool(acc), nd.value, fe.value


def dp45_run(p, psi, n, pump=None, pot=None, t=0.0, dt=0.04, t_max=1.0, tol=1e-6, dt_min=1e-4, dt_max=0.1,
             max_rejections=1000, psi_minus=None, n_minus=None, pump_minus=None, pot_minus=None):
    psi = _c128(psi)
    n = _res(p, n, psi.shape, keep_none=True)
    pm = _c128(psi_minus) if psi_minus is not None else None
    nm = _res(p, n_minus, psi.shape, keep_none=True)
    tt, dd = C.c_double(t), C.c_double(dt)
    nrej = C.c_int64()
    acc = lib().orc_dp45_run(C.byref(p), _dp(psi.view(np.float64)), _dp(pm.view(np.float64)) if pm is not None else None,
                             _rdp(n), _rdp(nm), _dp(_f64(pump)), _dp(_f64(pump_minus)), _dp(_f64(pot)),
                             _dp(_f64(pot_minus)), C.byref(tt), C.byref(dd), t_max, tol, dt_min, dt_max,
                             max_rejections, C.byref(nrej))
    if acc < 0:
        raise RuntimeError(f"orc_dp45_run failed: {acc}")
    res = dict(psi=psi, n=n, t=tt.value, dt=dd.value, accepted=int(acc), rejected=int(nrej.value))
    if pm is not None:
        res.update(psi_minus=pm, n_minus=nm)
    return res


ITERATORS = ("Newton", "MP", "Heun2", "Ralston", "Heun3", "Ralston3", "RK3", "SSPRK3", "HW", "RK4", "Ralston4", "Rule38",
             "Nystroem", "Fehlberg2", "BoSha", "CashKarp", "Fehlberg5", "DP45")


def iter_info(key):
    """(halo, adaptTimeStep power or 0, stages fixed, stages adaptive or 0) of an iterator key."""
    h, st_f, st_a = C.c_int(), C.c_int(), C.c_int()
    pw = C.c_double()
    if lib().orc_iter_info(key.encode(), C.byref(h), C.byref(pw), C.byref(st_f), C.byref(st_a)):
        raise KeyError(key)
    return h.value, pw.value, st_f.value, st_a.value


def iter_run(p, key, psi, n=None, pump=None, pot=None, dt=0.04, steps=1, psi_minus=None, n_minus=None, pump_minus=None,
             pot_minus=None, step0=0):
    """`steps` fixed steps of the iterator `key` (orc_iter_run); returns as rk4_run."""
    psi = _c128(psi)
    pm = _c128(psi_minus) if psi_minus is not None else None
    n = _res(p, n, psi.shape)
    nm = _res(p, n_minus, psi.shape) if (n_minus is not None or psi_minus is not None) else None
    rc = lib().orc_iter_run(C.byref(p), key.encode(), _dp(psi.view(np.float64)),
                            _dp(pm.view(np.float64)) if pm is not None else None, _rdp(n), _rdp(nm), _dp(_f64(pump)),
                            _dp(_f64(pump_minus)), _dp(_f64(pot)), _dp(_f64(pot_minus)), dt, steps, step0)
    if rc:
        raise RuntimeError(f"orc_iter_run({key}) failed: {rc}")
    if psi_minus is not None:
        return psi, n, pm, nm
    return psi, n


def iter_attempt(p, key, psi, n=None, pump=None, pot=None, dt=0.04, noise_dt=None, step=0, psi_minus=None, n_minus=None,
                 pump_minus=None, pot_minus=None):
    """One adaptive attempt of `key` (orc_iter_attempt); returns as dp45_attempt."""
    psi = _c128(psi)
    n = _res(p, n, psi.shape)
    pm = _c128(psi_minus) if psi_minus is not None else None
    nm = _res(p, n_minus, psi.shape) if (n_minus is not None or psi_minus is not None) else None
    es, nr = C.c_double(), C.c_double()
    rc = lib().orc_iter_attempt(C.byref(p), key.encode(), _dp(psi.view(np.float64)),
                                _dp(pm.view(np.float64)) if pm is not None else None, _rdp(n), _rdp(nm), _dp(_f64(pump)),
                                _dp(_f64(pump_minus)), _dp(_f64(pot)), _dp(_f64(pot_minus)), dt,
                                dt if noise_dt is None else noise_dt, step, C.byref(es), C.byref(nr))
    if rc:
        raise RuntimeError(f"orc_iter_attempt({key}) failed: {rc}")
    out = (psi, n, es.value, nr.value)
    if psi_minus is not None:
        out = out + (pm, nm)
    return out


def adapt_dt_pow(err_sum, norm, tol, dt, dt_min, dt_max, power):
    nd, fe = C.c_double(), C.c_double()
    acc = lib().orc_adapt_dt_pow(err_sum, norm, tol, dt, dt_min, dt_max, power, C.byref(nd), C.byref(fe))
    return bool(acc), nd.value, fe.value


def iter_adaptive_run(p, key, psi, n, pump=None, pot=None, t=0.0, dt=0.04, t_max=1.0, tol=1e-6, dt_min=1e-4, dt_max=0.1,
                      max_rejections=1000, psi_minus=None, n_minus=None, pump_minus=None, pot_minus=None):
    """Adaptive run of `key` until t >= t_max (orc_iter_adaptive_run); returns as dp45_run."""
    psi = _c128(psi)
    n = _res(p, n, psi.shape, keep_none=True)
    pm = _c128(psi_minus) if psi_minus is not None else None
    nm = _res(p, n_minus, psi.shape, keep_none=True)
    tt, dd = C.c_double(t), C.c_double(dt)
    nrej = C.c_int64()
    acc = lib().orc_iter_adaptive_run(C.byref(p), key.encode(), _dp(psi.view(np.float64)),
                                      _dp(pm.view(np.float64)) if pm is not None else None, _rdp(n), _rdp(nm),
                                      _dp(_f64(pump)), _dp(_f64(pump_minus)), _dp(_f64(pot)), _dp(_f64(pot_minus)),
                                      C.byref(tt), C.byref(dd), t_max, tol, dt_min, dt_max, max_rejections,
                                      C.byref(nrej))
    if acc < 0:
        raise RuntimeError(f"orc_iter_adaptive_run({key}) failed: {acc}")
    res = dict(psi=psi, n=n, t=tt.value, dt=dd.value, accepted=int(acc), rejected=int(nrej.value))
    if pm is not None:
        res.update(psi_minus=pm, n_minus=nm)
    return res


def ssfm_run(p, psi, n=None, pump=None, pot=None, dt=0.04, steps=1, psi_minus=None, n_minus=None, pump_minus=None,
             pot_minus=None, step0=0):
    """`steps` split-step Fourier steps (orc_ssfm_run); returns as rk4_run."""
    psi = _c128(psi)
    pm = _c128(psi_minus) if psi_minus is not None else None
    n = _res(p, n, psi.shape)
    nm = _res(p, n_minus, psi.shape) if (n_minus is not None or psi_minus is not None) else None
    rc = lib().orc_ssfm_run(C.byref(p), _dp(psi.view(np.float64)), _dp(pm.view(np.float64)) if pm is not None else None,
                            _rdp(n), _rdp(nm), _dp(_f64(pump)), _dp(_f64(pump_minus)), _dp(_f64(pot)), _dp(_f64(pot_minus)),
                            dt, steps, step0)
    if rc:
        raise RuntimeError(f"orc_ssfm_run failed: {rc}")
    if psi_minus is not None:
        return psi, n, pm, nm
    return psi, n


def fft_filter(psi, mask):
    psi = _c128(psi)
    mask = _f64(mask)
    Nr, Nc = psi.shape
    if lib().orc_fft_filter(_dp(psi.view(np.float64)), _dp(mask), Nc, Nr):
        raise RuntimeError("orc_fft_filter failed")
    return psi


def fft2d(psi, sign=-1):
    psi = _c128(psi)
    Nr, Nc = psi.shape
    lib().orc_fft2d(_dp(psi.view(np.float64)), Nc, Nr, sign)
    return psi


ENV_GAUSS, ENV_OUTER, ENV_RING, ENV_NODIVIDE, ENV_LOCAL = 1, 2, 4, 8, 16
BEH_ADD, BEH_MULTIPLY, BEH_REPLACE, BEH_ADAPTIVE, BEH_COMPLEX = 1, 2, 4, 8, 16


def envelope(N_c, N_r, L_x, L_y, dx, dy, amp, width_x, width_y, pos_x, pos_y, exponent, type_flags,
             charge=0.0, k0=(0.0, 0.0), behaviour=1, out=None):
    out = np.zeros((N_r, N_c), np.complex128) if out is None else _c128(out)
    lib().orc_envelope_add(_dp(out.view(np.float64)), N_c, N_r, L_x, L_y, dx, dy, amp, width_x, width_y, pos_x,
                           pos_y, exponent, type_flags, charge, k0[0], k0[1], behaviour)
    return out


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def noise_normal2(seed, step, cell, component=0):
    re, im = C.c_double(), C.c_double()
    lib().orc_noise_normal2(seed, step, cell, component, C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def gaussian_field(N_c, N_r, amplitude, seed):
    out = np.zeros((N_r, N_c), np.complex128)
    lib().orc_gaussian_field(_dp(out.view(np.float64)), N_c * N_r, amplitude, seed)
    return out


def mt19937_words(seed: int, count: int) -> np.ndarray:
    """The first `count` 32-bit outputs of std::mt19937{seed} (Matsumoto & Nishimura 1998, seeded by
    init_genrand, as the C++ standard specifies), from numpy's MT19937 with the same legacy seeding."""
    rs = np.random.RandomState(seed & 0xFFFFFFFF)
    return rs._bit_generator.random_raw(count).astype(np.uint64)


def init_random(N_c, N_r, amplitude, seed, fp32=False, components=1):
    """--initRandom amplitude seed (solver_initialization.cu:71-79): std::mt19937{seed} and
    std::uniform_real_distribution<real>(-a, a); each cell of psi+ (then psi-) in row-major order
    gets z += (u, u') with the real part drawn first.  The draws follow libstdc++'s
    generate_canonical: for float one 32-bit word w, u01 = float(w) / 2^32; for double two
    words, u01 = (w1 + w2 2^32) / 2^64 in double; u01 >= 1 becomes the largest value below 1;
    then u = u01 (b - a) + a in the real type.  Returns a list of complex128 (N_r, N_c) fields."""
    n = N_c * N_r
    draws = 2 * n * components
    per = 1 if fp32 else 2
    w = mt19937_words(seed, draws * per)
    if fp32:
        f = np.float32
        u01 = w.astype(np.float32) / f(4294967296.0)
        u01 = np.where(u01 >= f(1), np.nextafter(f(1), f(0)), u01).astype(np.float32)
        a, b = f(-amplitude), f(amplitude)
        u = (u01 * (b - a) + a).astype(np.float32).astype(np.float64)
    else:
        lo, hi = w[0::2].astype(np.float64), w[1::2].astype(np.float64)
        s = lo + hi * 4294967296.0
        u01 = s / 18446744073709551616.0
        u01 = np.where(u01 >= 1.0, np.nextafter(1.0, 0.0), u01)
        a, b = -float(amplitude), float(amplitude)
        u = u01 * (b - a) + a
    out = []
    for k in range(components):
        seg = u[2 * n * k:2 * n * (k + 1)]
        out.append((seg[0::2] + 1j * seg[1::2]).reshape(N_r, N_c))
    return out
