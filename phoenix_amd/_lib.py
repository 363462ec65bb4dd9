"""ctypes binding of libphx.so (the C ABI declared in include/phx.h).

The HIP library is mandatory: importing this module raises if libphx.so has not
been built (python -m phoenix_amd.build).  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libphx.so")

PHX_FP32, PHX_FP64 = 0, 1
PHX_OK, PHX_ERR_INVALID, PHX_ERR_HIP, PHX_ERR_UNSUPPORTED, PHX_ERR_IO, PHX_ERR_STEP, PHX_ERR_STATE = 0, -1, -2, -3, -4, -5, -6
FIELDS = {
    "psi_plus": 0, "psi_minus": 1, "n_plus": 2, "n_minus": 3, "pump_plus": 4, "pump_minus": 5,
    "potential_plus": 6, "potential_minus": 7, "fft_mask_plus": 8, "fft_mask_minus": 9,
    "fft_plus": 10, "fft_minus": 11,
}
COMPLEX_FIELDS = {0, 1, 10, 11}

# Every entry point declared in include/phx.h (checked by tests/test_capi.py).
EXPORTED = [
    "phx_create", "phx_initialize", "phx_step", "phx_run", "phx_apply_fft_filter", "phx_get_field", "phx_set_field",
    "phx_device_ptr", "phx_get_time", "phx_set_time", "phx_get_info", "phx_set_stream", "phx_sync", "phx_destroy",
    "phx_last_error", "phx_version", "phx_physics_defaults", "phx_rk4_scalar_device",
    "phx_ensemble_create", "phx_ensemble_set_field", "phx_ensemble_get_field", "phx_ensemble_device_ptr",
    "phx_ensemble_set_tile_rows", "phx_ensemble_step", "phx_ensemble_set_stream", "phx_ensemble_sync",
    "phx_ensemble_destroy",
    "phx_adjoint_create", "phx_adjoint_forward", "phx_adjoint_backward", "phx_adjoint_set_stream", "phx_adjoint_sync",
    "phx_adjoint_destroy",
    "phx_run_interval", "phx_source_groups", "phx_source_field", "phx_source_temporal", "phx_get_output_config",
    "phx_step_at", "phx_rk4_scalar_device_noalloc",
]
PHX_ENS_RESERVOIR, PHX_ENS_PUMP, PHX_ENS_POTENTIAL = 1, 2, 4
def stream_arg(stream_ptr: Optional[int]):
    """hipStream_t for the set_stream entry points, whose NULL selects the handle's own stream.
    None -> the handle's own stream.  A handle of 0 (torch's default stream) is refused: it would
    silently select the handle's own stream, unordered with the caller's work on the default stream;
    such callers run the handle on a stream of their own (torch.cuda.Stream) ordered by events."""
    if stream_ptr is None:
        return None
    if stream_ptr == 0:
        raise ValueError("stream handle 0 (the default stream) cannot be set on a handle: pass None for the "
                         "handle's own stream, or a torch.cuda.Stream ordered with the default stream")
    return stream_ptr


class PhxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[phx {code}] {msg}")
        self.code = code


class PhxInfo(C.Structure):
    _fields_ = [
        ("N_c", C.c_int32), ("N_r", C.c_int32), ("precision", C.c_int32), ("tetm", C.c_int32),
        ("use_reservoir", C.c_int32), ("use_pump", C.c_int32), ("use_potential", C.c_int32),
        ("use_stochastic", C.c_int32), ("use_fft_mask", C.c_int32), ("adaptive", C.c_int32),
        ("periodic_x", C.c_int32), ("periodic_y", C.c_int32), ("halo", C.c_int32), ("iterator", C.c_char * 16),
        ("dx", C.c_double), ("dy", C.c_double), ("dt", C.c_double), ("t", C.c_double), ("t_max", C.c_double),
        ("fft_every", C.c_double), ("tolerance", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double),
        ("iteration", C.c_int64), ("rejected", C.c_int64), ("tile_rows", C.c_int32), ("prefetch_rows", C.c_int32),
        ("staged", C.c_int32), ("multi_step_steps", C.c_int64), ("multi_step_launches", C.c_int64),
        ("complex_reservoir", C.c_int32),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["iterator"] = d["iterator"].decode()
        return d


class PhxPhysics(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("dx", "dy", "dV", "m_eff_scaled", "one_over_h_bar_s", "gamma_c", "gamma_r",
                                          "g_c", "g_r", "R", "g_pm", "delta_LT")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m phoenix_amd.build` (no CPU fallback)")
        # One HIP runtime per process: PyTorch ships its own libamdhip64 (SONAME libamdhip64.so.7) and
        # loads it by an unversioned name.  Loading torch first lets libphx.so bind to that same runtime;
        # loading libphx.so first would pull in a second copy and leave torch without devices.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        vp, i64 = C.c_void_p, C.c_int64
        L.phx_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.POINTER(C.c_char_p)]
        L.phx_initialize.argtypes = [vp]
        L.phx_step.argtypes = [vp, i64, C.POINTER(i64)]
        L.phx_run.argtypes = [vp, C.POINTER(i64)]
        L.phx_run_interval.argtypes = [vp, C.POINTER(i64), C.POINTER(C.c_int)]
        L.phx_step_at.argtypes = [vp, C.c_double, C.POINTER(C.c_double), C.c_int, C.POINTER(i64)]
        L.phx_apply_fft_filter.argtypes = [vp]
        L.phx_get_field.argtypes = [vp, C.c_int, vp, C.c_size_t]
        L.phx_set_field.argtypes = [vp, C.c_int, vp, C.c_size_t]
        L.phx_device_ptr.argtypes = [vp, C.c_int, C.POINTER(vp)]
        L.phx_get_time.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.phx_set_time.argtypes = [vp, C.c_double, C.c_double]
        L.phx_get_info.argtypes = [vp, C.POINTER(PhxInfo)]
        L.phx_set_stream.argtypes = [vp, vp]
        L.phx_sync.argtypes = [vp]
        L.phx_destroy.argtypes = [vp]
        L.phx_destroy.restype = None
        L.phx_last_error.restype = C.c_char_p
        L.phx_version.restype = C.c_char_p
        L.phx_physics_defaults.argtypes = [C.POINTER(PhxPhysics), C.c_int, C.c_int, C.c_double, C.c_double]
        L.phx_physics_defaults.restype = C.c_double
        L.phx_rk4_scalar_device.argtypes = [C.c_int, C.POINTER(PhxPhysics), C.c_int, C.c_int, C.c_int, C.c_int,
                                            vp, vp, vp, vp, vp, vp, C.c_double, i64, vp]
        L.phx_rk4_scalar_device_noalloc.argtypes = [C.c_int, C.POINTER(PhxPhysics), C.c_int, C.c_int, C.c_int, C.c_int,
                                                    vp, vp, vp, vp, C.c_double, i64, vp]
        ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.phx_ensemble_create.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(PhxPhysics), C.c_int, ip, ip, dp, dp,
                                          C.c_int, C.c_int, C.c_int]
        L.phx_ensemble_set_field.argtypes = [vp, C.c_int, C.c_int, vp, C.c_size_t]
        L.phx_ensemble_get_field.argtypes = [vp, C.c_int, C.c_int, vp, C.c_size_t]
        L.phx_ensemble_device_ptr.argtypes = [vp, C.c_int, C.c_int, C.POINTER(vp)]
        L.phx_ensemble_set_tile_rows.argtypes = [vp, C.c_int]
        L.phx_ensemble_step.argtypes = [vp, C.c_double, i64, C.c_int]
        L.phx_ensemble_set_stream.argtypes = [vp, vp]
        L.phx_ensemble_sync.argtypes = [vp]
        L.phx_ensemble_destroy.argtypes = [vp]
        L.phx_ensemble_destroy.restype = None
        L.phx_adjoint_create.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(PhxPhysics), C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.c_int, C.c_int, C.c_int, C.c_double, i64, C.c_int]
        L.phx_adjoint_forward.argtypes = [vp] * 7
        L.phx_adjoint_backward.argtypes = [vp] * 7
        L.phx_adjoint_set_stream.argtypes = [vp, vp]
        L.phx_adjoint_sync.argtypes = [vp]
        L.phx_adjoint_destroy.argtypes = [vp]
        L.phx_adjoint_destroy.restype = None
        _lib = L
    return _lib


def check(rc: int):
    if rc != PHX_OK:
        raise PhxError(rc, lib().phx_last_error().decode())
    return rc


class Solver:
    """One configured PHOENIX system on the GPU (handle of the C ABI).

    `args` are PHOENIX command line tokens, e.g. ["--N", "256", "256", "--tmax", "10"].
    """

    def __init__(self, args: Sequence[str], precision: str = "fp64", initialize: bool = True):
        self.precision = PHX_FP32 if precision == "fp32" else PHX_FP64
        self.real = np.float32 if self.precision == PHX_FP32 else np.float64
        self.cplx = np.complex64 if self.precision == PHX_FP32 else np.complex128
        argv = ["phoenix"] + [str(a) for a in args]
        arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
        h = C.c_void_p()
        check(lib().phx_create(C.byref(h), self.precision, len(argv), arr))
        self._h = h
        if initialize:
            self.initialize()

    def initialize(self):
        check(lib().phx_initialize(self._h))

    @property
    def info(self) -> dict:
        i = PhxInfo()
        check(lib().phx_get_info(self._h, C.byref(i)))
        return i.as_dict()

    def step(self, n: int = 1) -> int:
        done = C.c_int64()
        check(lib().phx_step(self._h, n, C.byref(done)))
        return done.value

    def run(self) -> int:
        done = C.c_int64()
        check(lib().phx_run(self._h, C.byref(done)))
        return done.value

    def run_interval(self):
        """One output interval of the main loop (phx_run_interval): (steps taken, finished)."""
        done, fin = C.c_int64(), C.c_int()
        check(lib().phx_run_interval(self._h, C.byref(done), C.byref(fin)))
        return done.value, bool(fin.value)

    def step_at(self, t: float, dt: float, variable_time_step: bool = False):
        """One Solver::step at kernel time t and step dt (phx_step_at): (dt for the next step,
        rejected attempts).  The handle's t is not advanced."""
        d, rej = C.c_double(dt), C.c_int64()
        check(lib().phx_step_at(self._h, t, C.byref(d), int(variable_time_step), C.byref(rej)))
        return d.value, rej.value

    def apply_fft_filter(self):
        check(lib().phx_apply_fft_filter(self._h))

    def _shape(self):
        i = self.info
        return i["N_r"], i["N_c"]

    def _dtype(self, fid: int):
        """Host dtype of a field: complex for psi and the spectra, and for the reservoir when it is complex"""
        cplx = fid in COMPLEX_FIELDS or (fid in (2, 3) and self.info["complex_reservoir"])
        return self.cplx if cplx else self.real

    def get(self, field: str) -> np.ndarray:
        fid = FIELDS[field]
        out = np.empty(self._shape(), self._dtype(fid))
        check(lib().phx_get_field(self._h, fid, out.ctypes.data, out.nbytes))
        return out

    def set(self, field: str, value: np.ndarray):
        fid = FIELDS[field]
        v = np.ascontiguousarray(value, self._dtype(fid))
        if v.shape != self._shape():
            raise ValueError(f"{field}: expected shape {self._shape()}, got {v.shape}")
        check(lib().phx_set_field(self._h, fid, v.ctypes.data, v.nbytes))

    def device_ptr(self, field: str) -> int:
        p = C.c_void_p()
        check(lib().phx_device_ptr(self._h, FIELDS[field], C.byref(p)))
        return p.value

    def time(self):
        t, dt = C.c_double(), C.c_double()
        check(lib().phx_get_time(self._h, C.byref(t), C.byref(dt)))
        return t.value, dt.value

    def set_time(self, t: float, dt: float):
        check(lib().phx_set_time(self._h, t, dt))

    def set_stream(self, stream_ptr: Optional[int]):
        check(lib().phx_set_stream(self._h, stream_arg(stream_ptr)))

    def sync(self):
        check(lib().phx_sync(self._h))

    def close(self):
        if getattr(self, "_h", None):
            lib().phx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def physics_defaults(N_c: int, N_r: int, L_x: float, L_y: float):
    ph = PhxPhysics()
    magic = lib().phx_physics_defaults(C.byref(ph), N_c, N_r, L_x, L_y)
    return ph, magic


def rk4_scalar_device(precision: int, ph: PhxPhysics, N_c: int, N_r: int, periodic: tuple, psi_a: int, n_a: int,
                      psi_b: int, n_b: int, pump: int, pot: int, dt: float, n_steps: int, stream: int = 0):
    check(lib().phx_rk4_scalar_device(precision, C.byref(ph), N_c, N_r, int(periodic[0]), int(periodic[1]), psi_a,
                                      n_a, psi_b, n_b, pump, pot, dt, n_steps, stream))


def rk4_scalar_device_noalloc(precision: int, ph: PhxPhysics, N_c: int, N_r: int, periodic: tuple, psi: int, n: int,
                              pump: int, pot: int, dt: float, n_steps: int, stream: Optional[int]):
    """phx_rk4_scalar_device_noalloc: `n_steps` RK4 steps in place on device pointers, with the
    library's fixed device scratch as the second buffer (no device allocation)."""
    check(lib().phx_rk4_scalar_device_noalloc(precision, C.byref(ph), N_c, N_r, int(periodic[0]), int(periodic[1]), psi, n,
                                              pump, pot, dt, n_steps, stream))


class Ensemble:
    """Independent scalar grids of different sizes stepped together by fixed RK4
    (phx_ensemble_* of include/phx.h): one fused launch per step for all of them.

    `sizes` is a list of (N_c, N_r); `spacing` an optional list of (dx, dy).
    """

    def __init__(self, precision: str, ph: PhxPhysics, sizes: Sequence[tuple], spacing: Optional[Sequence[tuple]] = None,
                 periodic=(True, True), reservoir: bool = True, pump: bool = True, potential: bool = False):
        self.precision = PHX_FP32 if precision == "fp32" else PHX_FP64
        self.real = np.float32 if self.precision == PHX_FP32 else np.float64
        self.cplx = np.complex64 if self.precision == PHX_FP32 else np.complex128
        self.sizes = [(int(a), int(b)) for a, b in sizes]
        n = len(self.sizes)
        Nc = (C.c_int * n)(*[a for a, _ in self.sizes])
        Nr = (C.c_int * n)(*[b for _, b in self.sizes])
        dx = dy = None
        if spacing is not None:
            dx = (C.c_double * n)(*[float(a) for a, _ in spacing])
            dy = (C.c_double * n)(*[float(b) for _, b in spacing])
        flags = (PHX_ENS_RESERVOIR if reservoir else 0) | (PHX_ENS_PUMP if pump else 0) | (PHX_ENS_POTENTIAL if potential else 0)
        h = C.c_void_p()
        check(lib().phx_ensemble_create(C.byref(h), self.precision, C.byref(ph), n, Nc, Nr, dx, dy, int(periodic[0]),
                                        int(periodic[1]), flags))
        self._h = h

    def __len__(self):
        return len(self.sizes)

    def _shape(self, i):
        return self.sizes[i][1], self.sizes[i][0]

    def get(self, i: int, field: str) -> np.ndarray:
        fid = FIELDS[field]
        out = np.empty(self._shape(i), self.cplx if fid in COMPLEX_FIELDS else self.real)
        check(lib().phx_ensemble_get_field(self._h, i, fid, out.ctypes.data, out.nbytes))
        return out

    def set(self, i: int, field: str, value: np.ndarray):
        fid = FIELDS[field]
        v = np.ascontiguousarray(value, self.cplx if fid in COMPLEX_FIELDS else self.real)
        if v.shape != self._shape(i):
            raise ValueError(f"{field}[{i}]: expected shape {self._shape(i)}, got {v.shape}")
        check(lib().phx_ensemble_set_field(self._h, i, fid, v.ctypes.data, v.nbytes))

    def device_ptr(self, i: int, field: str) -> int:
        p = C.c_void_p()
        check(lib().phx_ensemble_device_ptr(self._h, i, FIELDS[field], C.byref(p)))
        return p.value

    def set_tile_rows(self, ty: int):
        check(lib().phx_ensemble_set_tile_rows(self._h, ty))

    def step(self, dt: float, n: int = 1, graph_steps: int = 0):
        check(lib().phx_ensemble_step(self._h, dt, n, graph_steps))

    def set_stream(self, stream_ptr: Optional[int]):
        check(lib().phx_ensemble_set_stream(self._h, stream_arg(stream_ptr)))

    def sync(self):
        check(lib().phx_ensemble_sync(self._h))

    def close(self):
        if getattr(self, "_h", None):
            lib().phx_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Adjoint:
    """Checkpointed RK4 forward / discrete-adjoint backward of `batch` scalar grids of one size
    (phx_adjoint_* of include/phx.h).  Fields are device pointers (e.g. tensor.data_ptr()), batch-major:
    psi (batch, N_r, N_c) complex, n (batch, N_r, N_c) real; P and V (N_r, N_c) if shared_params,
    else (batch, N_r, N_c)."""

    def __init__(self, precision: str, ph: PhxPhysics, N_c: int, N_r: int, batch: int, dt: float, steps: int,
                 checkpoint_every: int = 50, periodic=(True, True), reservoir: bool = True, pump: bool = True,
                 potential: bool = True, shared_params: bool = True):
        self.precision = PHX_FP32 if precision == "fp32" else PHX_FP64
        flags = (PHX_ENS_RESERVOIR if reservoir else 0) | (PHX_ENS_PUMP if pump else 0) | (PHX_ENS_POTENTIAL if potential else 0)
        h = C.c_void_p()
        check(lib().phx_adjoint_create(C.byref(h), self.precision, C.byref(ph), N_c, N_r, batch, int(periodic[0]),
                                       int(periodic[1]), flags, int(shared_params), dt, steps, checkpoint_every))
        self._h = h

    def forward(self, psi0: int, n0: int, pump: int, pot: int, psi_T: int, n_T: int):
        check(lib().phx_adjoint_forward(self._h, psi0, n0, pump, pot, psi_T, n_T))

    def backward(self, g_psi_T: int, g_n_T: int, g_pump: int, g_pot: int, g_psi0: int, g_n0: int):
        check(lib().phx_adjoint_backward(self._h, g_psi_T, g_n_T, g_pump, g_pot, g_psi0, g_n0))

    def set_stream(self, stream_ptr: Optional[int]):
        check(lib().phx_adjoint_set_stream(self._h, stream_arg(stream_ptr)))

    def sync(self):
        check(lib().phx_adjoint_sync(self._h))

    def close(self):
        if getattr(self, "_h", None):
            lib().phx_adjoint_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
