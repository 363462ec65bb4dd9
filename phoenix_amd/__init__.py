"""phoenix_amd -- MI355X-native time stepping for PHOENIX (driven-dissipative GP + reservoir).

Public surface:
  Solver            in-process handle over the C ABI (include/phx.h)
  pyphoenix         drop-in for the reference's subprocess wrapper (pyphoenix/pyphoenix.py)
  build()           compile libphx.so and the PHOENIX_gpu_{fp32,fp64} executables
"""
from .build import build  # noqa: F401


def __getattr__(name):
    if name in ("Solver", "PhxError", "physics_defaults", "rk4_scalar_device", "lib"):
        from . import _lib
        return getattr(_lib, name)
    if name == "pyphoenix":
        from .pyphoenix import pyphoenix
        return pyphoenix
    raise AttributeError(name)
