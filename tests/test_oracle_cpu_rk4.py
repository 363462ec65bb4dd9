"""The CPU-baseline timing restatement (oracle/phx_cpu_rk4.c) computes the oracle's RK4 step.

Pins: its FP64 build equals orc_rk4_run bit for bit (same operation order, no contraction) for
every subgrid split, including the periodic halo of the padded subgrids; its FP32 build stays within
1e-5 (relative L2) of it.  So the timed CPU work is the reference's step, not a stand-in.
"""
import os

import numpy as np
import pytest

from oracle import cpu_bench as B
from oracle import oracle as O


def run(prec, N, split, steps, seed):
    L, psi, n, P = B.workload(N, True, seed)
    n = np.random.default_rng(seed).random((N, N))
    lib = B.load(prec, B.isa_level() if os.path.exists(os.path.join(B.HERE, f"libphx_cpu_rk4_{prec}_v4.so")) else "v3")
    B.setup(lib, N, split, psi, n, P, 0.04, B.constants(N, L))
    lib.cpu_rk4_run(steps, None)
    po = np.zeros((N, N), np.complex128)
    no = np.zeros((N, N))
    import ctypes as C
    lib.cpu_rk4_get(po.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), no.ctypes.data_as(C.POINTER(C.c_double)))
    lib.cpu_rk4_free()
    p = O.params(N, N, L, L, pump=True)
    o, on = O.rk4_run(p, psi, n, pump=P, dt=0.04, steps=steps)
    return po, no, o, on


@pytest.mark.parametrize("split", [1, 2, 4, 8])
def test_fp64_equals_oracle_bitwise(split):
    po, no, o, on = run("fp64", 64, split, 7, 3)
    assert np.array_equal(po, o) and np.array_equal(no, on)


def test_fp32_close_to_oracle():
    po, no, o, on = run("fp32", 64, 4, 50, 5)
    assert np.linalg.norm(po - o) / np.linalg.norm(o) <= 1e-5
    assert np.linalg.norm(no - on) / np.linalg.norm(on) <= 1e-5
