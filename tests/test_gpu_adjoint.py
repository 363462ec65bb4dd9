"""GPU tests of the discrete adjoint (phx_adjoint_*, rk4_adjoint_fused) and of the
torch.autograd.Function over it (phoenix_amd.adjoint.rk4_evolve).

Bars (stated per test):
  - FP64 gradients (dJ/dP, dJ/dV, dJ/dpsi0, dJ/dn0) vs the FP64 CPU adjoint of the
    oracle: rel L2 <= 1e-10;
  - FP64 gradients vs central finite differences of the GPU forward at 64 random
    pixels: relative 1e-4 where the difference quotient is well conditioned;
  - FP32 gradients vs FP64 gradients: rel L2 <= 1e-3;
  - forward result bitwise equal to the plain RK4 entry point; gradients bitwise
    independent of the checkpoint spacing and of repeated backward passes.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

PTH = 0.15 * 0.225 / 0.01


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def problem(N_c, N_r, L, B, seed=5):
    dx, dy = L / (N_c - 1), L / (N_r - 1)
    x = -L / 2 + dx * np.arange(N_c)
    y = -L / 2 + dy * np.arange(N_r)
    X, Y = np.meshgrid(x, y)
    P = 2.0 * PTH * np.exp(-(X ** 2 + Y ** 2) / (2 * 6.0 ** 2))
    V = 3e-4 * np.exp(-((X - 3) ** 2 + (Y + 2) ** 2) / (2 * 4.0 ** 2))
    rng = np.random.default_rng(seed)
    psi0 = 0.2 * (rng.standard_normal((B, N_r, N_c)) + 1j * rng.standard_normal((B, N_r, N_c)))
    n0 = rng.random((B, N_r, N_c))
    mask = ((X - 4) ** 2 + (Y - 3) ** 2 < 5.0 ** 2).astype(float)
    return P, V, psi0, n0, mask


def torch_run(prec, N_c, N_r, L, B, steps, dt, K, periodic, P, V, psi0, n0, mask, need=("P", "V", "psi", "n")):
    import torch
    from phoenix_amd import _lib
    from phoenix_amd.adjoint import rk4_evolve
    ph, _ = _lib.physics_defaults(N_c, N_r, L, L)
    cd = torch.complex64 if prec == "fp32" else torch.complex128
    rd = torch.float32 if prec == "fp32" else torch.float64
    dev = torch.device("cuda:0")
    tP = torch.tensor(P, dtype=rd, device=dev, requires_grad="P" in need)
    tV = torch.tensor(V, dtype=rd, device=dev, requires_grad="V" in need)
    tpsi = torch.tensor(psi0, dtype=cd, device=dev, requires_grad="psi" in need)
    tn = torch.tensor(n0, dtype=rd, device=dev, requires_grad="n" in need)
    tm = torch.tensor(mask, dtype=rd, device=dev)
    psiT, nT = rk4_evolve(tP, tV, tpsi, tn, physics=ph, dt=dt, steps=steps, checkpoint_every=K,
                          periodic=(periodic, periodic))
    J = (psiT.abs() ** 2 * tm).sum() + 0.3 * (nT * tm).sum()
    J.backward()
    torch.cuda.synchronize()
    g = lambda t: None if t.grad is None else t.grad.detach().cpu().numpy()
    return float(J), g(tP), g(tV), g(tpsi), g(tn), psiT.detach().cpu().numpy(), nT.detach().cpu().numpy()


def oracle_grads(N_c, N_r, L, steps, dt, periodic, P, V, psi0, n0, mask):
    p = O.params(N_c, N_r, L, L, periodic=(periodic, periodic), reservoir=True, pump=True, potential=True)
    gP, gV, gpsi, gn = np.zeros_like(P), np.zeros_like(V), np.zeros_like(psi0), np.zeros_like(n0)
    for b in range(psi0.shape[0]):
        psiT, nT = O.rk4_run(p, psi0[b], n0[b], pump=P, pot=V, dt=dt, steps=steps)
        lp, ln, dP, dV, *_ = O.rk4_adjoint(p, psi0[b], n0[b], P, V, dt, steps, 2 * mask * psiT, 0.3 * mask)
        gP += dP
        gV += dV
        gpsi[b] = lp
        gn[b] = ln
    return gP, gV, gpsi, gn


CASES = [
    # (N_c, N_r, L, batch, steps, K, periodic)
    (48, 40, 23.5, 2, 37, 10, True),
    (60, 52, 29.5, 1, 30, 7, False),
    (130, 70, 64.5, 3, 24, 50, True),  # three strips, one checkpoint interval
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}x{c[1]}_b{c[3]}_{'per' if c[6] else 'zero'}" for c in CASES])
def test_adjoint_fp64_matches_oracle_adjoint(case):
    N_c, N_r, L, B, steps, K, periodic = case
    say(f"[adjoint vs oracle] case={case}")
    P, V, psi0, n0, mask = problem(N_c, N_r, L, B)
    _, gP, gV, gpsi, gn, _, _ = torch_run("fp64", N_c, N_r, L, B, steps, 0.02, K, periodic, P, V, psi0, n0, mask)
    oP, oV, opsi, on = oracle_grads(N_c, N_r, L, steps, 0.02, periodic, P, V, psi0, n0, mask)
    errs = [rel_l2(gP, oP), rel_l2(gV, oV), rel_l2(gpsi, opsi), rel_l2(gn, on)]
    say(f"  rel L2 dP {errs[0]:.2e} dV {errs[1]:.2e} dpsi0 {errs[2]:.2e} dn0 {errs[3]:.2e}")
    assert max(errs) <= 1e-10


def test_adjoint_fp64_matches_finite_differences():
    import torch
    from phoenix_amd import _lib
    N_c, N_r, L, B, steps, dt, K = 40, 40, 19.5, 1, 40, 0.02, 9
    say("[adjoint vs FD] 40x40, 64 pixels")
    P, V, psi0, n0, mask = problem(N_c, N_r, L, B, seed=11)
    _, gP, gV, _, _, _, _ = torch_run("fp64", N_c, N_r, L, B, steps, dt, K, True, P, V, psi0, n0, mask, need=("P", "V"))
    ph, _ = _lib.physics_defaults(N_c, N_r, L, L)
    dev = torch.device("cuda:0")
    tpsi = torch.tensor(psi0[0], dtype=torch.complex128, device=dev)
    tn = torch.tensor(n0[0], dtype=torch.float64, device=dev)
    tm = torch.tensor(mask, dtype=torch.float64, device=dev)

    def J(Pv, Vv):
        a, b = tpsi.clone(), tn.clone()
        a2, b2 = torch.empty_like(a), torch.empty_like(b)
        tP = torch.tensor(Pv, dtype=torch.float64, device=dev)
        tV = torch.tensor(Vv, dtype=torch.float64, device=dev)
        _lib.rk4_scalar_device(_lib.PHX_FP64, ph, N_c, N_r, (True, True), a.data_ptr(), b.data_ptr(), a2.data_ptr(),
                               b2.data_ptr(), tP.data_ptr(), tV.data_ptr(), dt, steps,
                               torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        return float((a.abs() ** 2 * tm).sum() + 0.3 * (b * tm).sum())

    rng = np.random.default_rng(2024)
    pix = [tuple(rng.integers(0, [N_r, N_c])) for _ in range(64)]
    worst, used = 0.0, 0
    for (r, c) in pix:
        for g, base, h, is_P in ((gP, P, 1e-4 * PTH, True), (gV, V, 1e-8, False)):
            a, b = base.copy(), base.copy()
            a[r, c] += h
            b[r, c] -= h
            fd = ((J(a, V) - J(b, V)) if is_P else (J(P, a) - J(P, b))) / (2 * h)
            if abs(fd) > 1e-3 * np.abs(g).max():
                worst = max(worst, abs(g[r, c] - fd) / abs(fd))
                used += 1
    say(f"  {used} well-conditioned differences, worst relative error {worst:.2e}")
    assert used >= 32
    assert worst <= 1e-4


def test_adjoint_fp32_close_to_fp64():
    N_c, N_r, L, B, steps, K = 64, 64, 31.5, 2, 60, 20
    say("[adjoint fp32 vs fp64] 64x64")
    P, V, psi0, n0, mask = problem(N_c, N_r, L, B)
    d = torch_run("fp64", N_c, N_r, L, B, steps, 0.02, K, True, P, V, psi0, n0, mask)
    s = torch_run("fp32", N_c, N_r, L, B, steps, 0.02, K, True, P, V, psi0, n0, mask)
    errs = [rel_l2(s[k], d[k]) for k in (1, 2, 3, 4)]
    say(f"  rel L2 fp32 vs fp64: {errs}")
    assert max(errs) <= 1e-3


def test_adjoint_forward_and_checkpoint_invariance():
    import torch
    from phoenix_amd import _lib
    N_c, N_r, L, B, steps, dt = 72, 56, 35.5, 2, 45, 0.02
    say("[adjoint invariance] 72x56")
    P, V, psi0, n0, mask = problem(N_c, N_r, L, B)
    a = torch_run("fp32", N_c, N_r, L, B, steps, dt, 7, True, P, V, psi0, n0, mask)
    b = torch_run("fp32", N_c, N_r, L, B, steps, dt, 45, True, P, V, psi0, n0, mask)
    for k in (1, 2, 3, 4, 5, 6):
        assert np.array_equal(a[k], b[k])
    # forward result = plain RK4 entry point, bit for bit
    ph, _ = _lib.physics_defaults(N_c, N_r, L, L)
    dev = torch.device("cuda:0")
    for i in range(B):
        tpsi = torch.tensor(psi0[i], dtype=torch.complex64, device=dev)
        tn = torch.tensor(n0[i], dtype=torch.float32, device=dev)
        tP = torch.tensor(P, dtype=torch.float32, device=dev)
        tV = torch.tensor(V, dtype=torch.float32, device=dev)
        a2, b2 = torch.empty_like(tpsi), torch.empty_like(tn)
        _lib.rk4_scalar_device(_lib.PHX_FP32, ph, N_c, N_r, (True, True), tpsi.data_ptr(), tn.data_ptr(), a2.data_ptr(),
                               b2.data_ptr(), tP.data_ptr(), tV.data_ptr(), dt, steps,
                               torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        assert np.array_equal(tpsi.cpu().numpy(), a[5][i])
        assert np.array_equal(tn.cpu().numpy(), a[6][i])


def test_adjoint_interleaved_forward_passes():
    """Two forward passes share one checkpoint store; the first one's backward rebuilds its checkpoints."""
    import torch
    from phoenix_amd import _lib
    from phoenix_amd.adjoint import rk4_evolve
    N_c, N_r, L, steps, dt = 48, 48, 23.5, 30, 0.02
    say("[adjoint interleaved]")
    P, V, psi0, n0, mask = problem(N_c, N_r, L, 1)
    ref = torch_run("fp64", N_c, N_r, L, 1, steps, dt, 10, True, P, V, psi0, n0, mask, need=("P",))[1]
    ph, _ = _lib.physics_defaults(N_c, N_r, L, L)
    dev = torch.device("cuda:0")
    tP = torch.tensor(P, dtype=torch.float64, device=dev, requires_grad=True)
    tV = torch.tensor(V, dtype=torch.float64, device=dev)
    tm = torch.tensor(mask, dtype=torch.float64, device=dev)
    psi_a = torch.tensor(psi0, dtype=torch.complex128, device=dev)
    n_a = torch.tensor(n0, dtype=torch.float64, device=dev)
    x1, m1 = rk4_evolve(tP, tV, psi_a, n_a, physics=ph, dt=dt, steps=steps, checkpoint_every=10)
    x2, m2 = rk4_evolve(tP, tV, 0.5 * psi_a, n_a, physics=ph, dt=dt, steps=steps, checkpoint_every=10)
    J1 = (x1.abs() ** 2 * tm).sum() + 0.3 * (m1 * tm).sum()
    J1.backward(retain_graph=True)
    g1 = tP.grad.detach().cpu().numpy().copy()
    tP.grad = None
    J1.backward()
    g1b = tP.grad.detach().cpu().numpy()
    assert np.array_equal(g1, ref)
    assert np.array_equal(g1b, ref)
    del x2, m2


def test_adjoint_per_item_parameters():
    """P and V with a batch dimension: each item gets its own gradient (FP64 vs the oracle, 1e-10)."""
    import torch
    from phoenix_amd import _lib
    from phoenix_amd.adjoint import rk4_evolve
    N_c, N_r, L, B, steps, dt = 40, 48, 23.5, 2, 25, 0.02
    say("[adjoint per-item P, V]")
    P, V, psi0, n0, mask = problem(N_c, N_r, L, B, seed=9)
    Pb = np.stack([P, 0.7 * P[::-1, :]])
    Vb = np.stack([V, -V[:, ::-1]])
    ph, _ = _lib.physics_defaults(N_c, N_r, L, L)
    dev = torch.device("cuda:0")
    tP = torch.tensor(Pb, dtype=torch.float64, device=dev, requires_grad=True)
    tV = torch.tensor(Vb, dtype=torch.float64, device=dev, requires_grad=True)
    tpsi = torch.tensor(psi0, dtype=torch.complex128, device=dev)
    tn = torch.tensor(n0, dtype=torch.float64, device=dev)
    tm = torch.tensor(mask, dtype=torch.float64, device=dev)
    psiT, nT = rk4_evolve(tP, tV, tpsi, tn, physics=ph, dt=dt, steps=steps, checkpoint_every=8)
    ((psiT.abs() ** 2 * tm).sum() + 0.3 * (nT * tm).sum()).backward()
    p = O.params(N_c, N_r, L, L, periodic=(True, True), reservoir=True, pump=True, potential=True)
    for b in range(B):
        oT, onT = O.rk4_run(p, psi0[b], n0[b], pump=Pb[b], pot=Vb[b], dt=dt, steps=steps)
        _, _, dP, dV, *_ = O.rk4_adjoint(p, psi0[b], n0[b], Pb[b], Vb[b], dt, steps, 2 * mask * oT, 0.3 * mask)
        e = max(rel_l2(tP.grad[b].cpu().numpy(), dP), rel_l2(tV.grad[b].cpu().numpy(), dV))
        say(f"  item {b}: rel L2 {e:.2e}")
        assert e <= 1e-10


def test_adjoint_without_reservoir_gradient_of_potential():
    """Reservoir off (psi-only model with a potential) through the C ABI: dJ/dV vs the oracle."""
    import torch
    from phoenix_amd import _lib
    N, L, steps, dt = 48, 23.5, 30, 0.02
    say("[adjoint no reservoir]")
    _, V, psi0, _, mask = problem(N, N, L, 1, seed=4)
    ph, _ = _lib.physics_defaults(N, N, L, L)
    dev = torch.device("cuda:0")
    tV = torch.tensor(V, dtype=torch.float64, device=dev)
    tpsi = torch.tensor(psi0, dtype=torch.complex128, device=dev)
    psiT = torch.empty_like(tpsi)
    eng = _lib.Adjoint("fp64", ph, N, N, 1, dt, steps, 7, (True, True), reservoir=False, pump=False, potential=True)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())  # the engine's stream after the tensors' initialisation
    eng.set_stream(side.cuda_stream)
    eng.forward(tpsi.data_ptr(), None, None, tV.data_ptr(), psiT.data_ptr(), None)
    torch.cuda.current_stream().wait_stream(side)
    g = (torch.tensor(2 * mask, dtype=torch.float64, device=dev) * psiT).contiguous()
    gV = torch.empty_like(tV)
    side.wait_stream(torch.cuda.current_stream())
    eng.backward(g.data_ptr(), None, None, gV.data_ptr(), None, None)
    torch.cuda.synchronize()
    p = O.params(N, N, L, L, periodic=(True, True), reservoir=False, pump=False, potential=True)
    oT, _ = O.rk4_run(p, psi0[0], None, pot=V, dt=dt, steps=steps)
    assert rel_l2(psiT[0].cpu().numpy(), oT) <= 1e-12
    _, _, _, dV, *_ = O.rk4_adjoint(p, psi0[0], None, None, V, dt, steps, 2 * mask * oT)
    e = rel_l2(gV.cpu().numpy(), dV)
    say(f"  rel L2 dV {e:.2e}")
    assert e <= 1e-10
    eng.close()


def test_default_stream_ordering():
    """rk4_evolve on torch's default stream (handle 0): the engine's work is ordered with torch's own
    kernels on that stream (the asynchronous zero fill of n0 before, the objective's reduction after),
    so the objective equals the one computed from the plain entry point's result."""
    import torch
    from phoenix_amd import _lib
    from phoenix_amd.adjoint import release_engines, rk4_evolve
    N, L, B, steps, dt = 256, 127.5, 8, 400, 0.04
    say("[adjoint default stream] 8 x 256^2, 400 steps")
    P, V, psi0, _, mask = problem(N, N, L, B, seed=3)
    dev = torch.device("cuda:0")
    assert torch.cuda.current_stream().cuda_stream == 0
    ph, _ = _lib.physics_defaults(N, N, L, L)
    tm = torch.tensor(mask, dtype=torch.float64, device=dev)
    tP = torch.tensor(P, dtype=torch.float64, device=dev, requires_grad=True)
    tV = torch.tensor(V, dtype=torch.float64, device=dev)
    tpsi = torch.tensor(psi0, dtype=torch.complex128, device=dev)
    tn = torch.zeros(psi0.shape, dtype=torch.float64, device=dev)  # asynchronous fill on the default stream
    psiT, _ = rk4_evolve(tP, tV, tpsi, tn, physics=ph, dt=dt, steps=steps, checkpoint_every=50)
    J = float((psiT.abs() ** 2 * tm).sum())  # reduction queued right behind the engine's last launch
    release_engines()
    ref = 0.0
    for b in range(B):
        x = torch.tensor(psi0[b], dtype=torch.complex128, device=dev)
        n = torch.zeros((N, N), dtype=torch.float64, device=dev)
        x2, n2 = torch.empty_like(x), torch.empty_like(n)
        _lib.rk4_scalar_device(_lib.PHX_FP64, ph, N, N, (True, True), x.data_ptr(), n.data_ptr(), x2.data_ptr(),
                               n2.data_ptr(), tP.detach().data_ptr(), tV.data_ptr(), dt, steps, None)
        torch.cuda.synchronize()
        ref += float((np.abs(x.cpu().numpy()) ** 2 * mask).sum())
    say(f"  J {J:.15e} reference {ref:.15e}")
    assert abs(J - ref) <= 1e-12 * abs(ref)
