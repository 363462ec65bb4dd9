"""GPU test of phx_rk4_scalar_device_noalloc (no device allocation; the double buffer is the
library's fixed device array) against phx_rk4_scalar_device, through the C ABI.

Bar: bitwise equal results (the same kernels run; only the second buffer differs), on two calls
in a row on the same device, the second with a smaller grid, for FP32 and FP64, odd and even
step counts (an odd count copies the result back from the array); a grid larger than the array
is rejected with PHX_ERR_INVALID before any device work.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
PTH = 0.15 * 0.225 / 0.01


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def inputs(N_c, N_r, seed):
    L = 0.5 * (N_c - 1)
    x = -L / 2 + 0.5 * np.arange(N_c)
    y = -0.5 * (N_r - 1) / 2 + 0.5 * np.arange(N_r)
    X, Y = np.meshgrid(x, y)
    rng = np.random.default_rng(seed)
    psi = 0.01 * (rng.standard_normal((N_r, N_c)) + 1j * rng.standard_normal((N_r, N_c)))
    n = rng.random((N_r, N_c))
    P = 1.5 * PTH * np.exp(-(X ** 2 + Y ** 2) / (2 * 8.0 ** 2))
    V = 2e-4 * np.exp(-((X - 3) ** 2 + Y ** 2) / (2 * 5.0 ** 2))
    return L, psi, n, P, V


def run_pair(prec, N_c, N_r, steps, seed, with_pot):
    import torch
    from phoenix_amd import _lib
    L, psi, n, P, V = inputs(N_c, N_r, seed)
    ph, _ = _lib.physics_defaults(N_c, N_r, L, 0.5 * (N_r - 1))
    cd = torch.complex64 if prec == "fp32" else torch.complex128
    rd = torch.float32 if prec == "fp32" else torch.float64
    code = _lib.PHX_FP32 if prec == "fp32" else _lib.PHX_FP64
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    tP = torch.tensor(P, dtype=rd, device=dev)
    tV = torch.tensor(V, dtype=rd, device=dev) if with_pot else None
    a_psi, a_n = torch.tensor(psi, dtype=cd, device=dev), torch.tensor(n, dtype=rd, device=dev)
    b_psi, b_n = torch.empty_like(a_psi), torch.empty_like(a_n)
    _lib.rk4_scalar_device(code, ph, N_c, N_r, (True, True), a_psi.data_ptr(), a_n.data_ptr(), b_psi.data_ptr(),
                           b_n.data_ptr(), tP.data_ptr(), tV.data_ptr() if with_pot else None, 0.04, steps, stream)
    c_psi, c_n = torch.tensor(psi, dtype=cd, device=dev), torch.tensor(n, dtype=rd, device=dev)
    _lib.rk4_scalar_device_noalloc(code, ph, N_c, N_r, (True, True), c_psi.data_ptr(), c_n.data_ptr(), tP.data_ptr(),
                                   tV.data_ptr() if with_pot else None, 0.04, steps, stream)
    torch.cuda.synchronize()
    return (a_psi.cpu().numpy(), a_n.cpu().numpy()), (c_psi.cpu().numpy(), c_n.cpu().numpy()), psi


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_noalloc_equals_device_entry(prec):
    cases = [(1024, 1024, 31, 3, False), (200, 136, 24, 4, True)] if prec == "fp32" else \
        [(1000, 1000, 17, 5, True), (96, 160, 30, 6, False)]
    for N_c, N_r, steps, seed, pot in cases:  # the second call is smaller than the first
        say(f"[noalloc] prec={prec} grid={N_c}x{N_r} steps={steps} pot={pot}")
        ref, got, psi0 = run_pair(prec, N_c, N_r, steps, seed, pot)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
        assert np.abs(ref[0] - psi0).max() > 1e-6  # the state moved


def test_noalloc_rejects_grids_beyond_the_array():
    import torch
    from phoenix_amd import _lib
    ph, _ = _lib.physics_defaults(2048, 2048, 1023.5, 1023.5)
    for prec, N, cd, rd in ((_lib.PHX_FP32, 2048, torch.complex64, torch.float32),
                            (_lib.PHX_FP64, 1200, torch.complex128, torch.float64)):
        psi = torch.zeros((N, N), dtype=cd, device="cuda")
        n = torch.zeros((N, N), dtype=rd, device="cuda")
        before = psi.clone()
        with pytest.raises(_lib.PhxError) as e:
            _lib.rk4_scalar_device_noalloc(prec, ph, N, N, (True, True), psi.data_ptr(), n.data_ptr(), None, None, 0.04, 1,
                                           None)
        assert e.value.code == _lib.PHX_ERR_INVALID
        torch.cuda.synchronize()
        assert torch.equal(psi, before)
