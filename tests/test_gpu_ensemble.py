"""GPU tests of the batched ensemble (phx_ensemble_*, rk4_batched_fused).

Bars:
  - every instance of an ensemble equals the same grid stepped alone through
    phx_rk4_scalar_device, bit for bit (same arithmetic per cell);
  - graph replay equals step-by-step launches bit for bit, also when the step
    count is not a multiple of the captured length;
  - FP64 instances match the FP64 oracle to rel L2 <= 1e-12 after 60 steps.
Instances have different sizes, spacings, pumps and initial states.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

PTH = 0.15 * 0.225 / 0.01

# (N_c, N_r, L): ragged sizes around the 56-column strip and the tile heights
# (spacings 0.40 .. 1.11, all stable at dt = 0.02)
SIZES = [(64, 64, 31.5), (100, 37, 40.0), (57, 130, 60.0), (128, 96, 70.0), (8, 8, 3.5), (200, 150, 99.5)]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def inputs(i, N_c, N_r, L):
    dx, dy = L / (N_c - 1), L / (N_r - 1)
    psi = O.gaussian_field(N_c, N_r, 0.01, 500 + i)
    x = -L / 2 + dx * np.arange(N_c)
    y = -L / 2 + dy * np.arange(N_r)
    X, Y = np.meshgrid(x, y)
    P = (0.8 + 0.4 * i) * PTH * np.exp(-0.5 * (X ** 2 + Y ** 2) / 8.0 ** 2)
    n = 0.5 * np.random.default_rng(i).random((N_r, N_c))
    return psi, n, P, dx, dy


def make_ensemble(prec, periodic=True):
    from phoenix_amd import _lib as L
    ph, _ = L.physics_defaults(64, 64, 31.5, 31.5)
    spacing, data = [], []
    for i, (N_c, N_r, Lg) in enumerate(SIZES):
        psi, n, P, dx, dy = inputs(i, N_c, N_r, Lg)
        spacing.append((dx, dy))
        data.append((psi, n, P))
    e = L.Ensemble(prec, ph, [(a, b) for a, b, _ in SIZES], spacing, periodic=(periodic, periodic))
    for i, (psi, n, P) in enumerate(data):
        e.set(i, "psi_plus", psi)
        e.set(i, "n_plus", n)
        e.set(i, "pump_plus", P)
    return e, ph, spacing, data


def alone(prec, ph, spacing_i, size_i, data_i, steps, dt, periodic):
    """One instance stepped through phx_rk4_scalar_device on torch buffers."""
    import torch
    from phoenix_amd import _lib as L
    N_c, N_r, _ = size_i
    psi, n, P = data_i
    cd = torch.complex64 if prec == "fp32" else torch.complex128
    rd = torch.float32 if prec == "fp32" else torch.float64
    dev = torch.device("cuda:0")
    tpsi = torch.from_numpy(psi.copy()).to(dev, cd)
    tn = torch.from_numpy(n.copy()).to(dev, rd)
    tP = torch.from_numpy(P.copy()).to(dev, rd)
    bpsi, bn = torch.empty_like(tpsi), torch.empty_like(tn)
    phi = L.PhxPhysics.from_buffer_copy(ph)
    phi.dx, phi.dy = spacing_i
    phi.dV = spacing_i[0] * spacing_i[1]
    L.rk4_scalar_device(L.PHX_FP32 if prec == "fp32" else L.PHX_FP64, phi, N_c, N_r, (periodic, periodic),
                        tpsi.data_ptr(), tn.data_ptr(), bpsi.data_ptr(), bn.data_ptr(), tP.data_ptr(), None, dt, steps,
                        torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return tpsi.cpu().numpy(), tn.cpu().numpy()


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("periodic", [True, False])
def test_ensemble_equals_instances_alone(prec, periodic):
    steps, dt = 33, 0.02
    say(f"[ensemble vs alone] prec={prec} periodic={periodic} steps={steps}")
    e, ph, spacing, data = make_ensemble(prec, periodic)
    e.step(dt, steps)
    e.sync()
    for i, size in enumerate(SIZES):
        g_psi, g_n = e.get(i, "psi_plus"), e.get(i, "n_plus")
        a_psi, a_n = alone(prec, ph, spacing[i], size, data[i], steps, dt, periodic)
        say(f"  instance {i} {size[:2]}: max |d psi| = {np.abs(g_psi - a_psi).max():.3e}")
        assert np.array_equal(g_psi, a_psi)
        assert np.array_equal(g_n, a_n)
    e.close()


@pytest.mark.parametrize("ty", [8, 10, 12, 16, 20])
def test_ensemble_graph_and_tile_rows(ty):
    steps, dt = 23, 0.02
    say(f"[ensemble graph] ty={ty} steps={steps}")
    a, *_ = make_ensemble("fp32")
    b, *_ = make_ensemble("fp32")
    a.set_tile_rows(12)
    b.set_tile_rows(ty)
    a.step(dt, steps)
    b.step(dt, steps, graph_steps=6)
    a.sync()
    b.sync()
    for i in range(len(SIZES)):
        assert np.array_equal(a.get(i, "psi_plus"), b.get(i, "psi_plus"))
        assert np.array_equal(a.get(i, "n_plus"), b.get(i, "n_plus"))
    # a second call replays the cached graph from the other buffer parity
    a.step(dt, 9)
    b.step(dt, 9, graph_steps=6)
    a.sync()
    b.sync()
    for i in range(len(SIZES)):
        assert np.array_equal(a.get(i, "psi_plus"), b.get(i, "psi_plus"))
    a.close()
    b.close()


def test_ensemble_fp64_matches_oracle():
    steps, dt = 60, 0.02
    say(f"[ensemble vs oracle] fp64 steps={steps}")
    e, ph, spacing, data = make_ensemble("fp64")
    e.step(dt, steps, graph_steps=10)
    e.sync()
    for i, (N_c, N_r, Lg) in enumerate(SIZES):
        psi, n, P = data[i]
        p = O.params(N_c, N_r, Lg, Lg, periodic=(True, True), reservoir=True, pump=True)
        o_psi, o_n = O.rk4_run(p, psi, n, pump=P, dt=dt, steps=steps)
        ep, en = rel_l2(e.get(i, "psi_plus"), o_psi), rel_l2(e.get(i, "n_plus"), o_n)
        say(f"  instance {i} ({N_c}x{N_r}): rel L2 psi {ep:.3e}  n {en:.3e}")
        assert ep <= 1e-12 and en <= 1e-12
    e.close()


@pytest.mark.parametrize("flags", [dict(reservoir=False, pump=False, potential=True),
                                   dict(reservoir=True, pump=True, potential=True)])
def test_ensemble_feature_sets_match_oracle(flags):
    """Ensembles without a reservoir (psi only, potential) and with every field, FP64 vs the oracle."""
    from phoenix_amd import _lib as L
    steps, dt = 30, 0.02
    say(f"[ensemble features] {flags}")
    ph, _ = L.physics_defaults(64, 64, 31.5, 31.5)
    sizes = [(48, 40, 23.5), (72, 64, 35.5)]
    e = L.Ensemble("fp64", ph, [(a, b) for a, b, _ in sizes], [(Lg / (a - 1), Lg / (b - 1)) for a, b, Lg in sizes], **flags)
    data = []
    for i, (N_c, N_r, Lg) in enumerate(sizes):
        psi, n, P, dx, dy = inputs(i, N_c, N_r, Lg)
        x = -Lg / 2 + dx * np.arange(N_c)
        y = -Lg / 2 + dy * np.arange(N_r)
        X, Y = np.meshgrid(x, y)
        V = 3e-4 * np.exp(-((X - 2) ** 2 + (Y + 1) ** 2) / 20.0)
        e.set(i, "psi_plus", psi)
        if flags["reservoir"]:
            e.set(i, "n_plus", n)
            e.set(i, "pump_plus", P)
        e.set(i, "potential_plus", V)
        data.append((psi, n, P, V))
    e.step(dt, steps)
    e.sync()
    for i, (N_c, N_r, Lg) in enumerate(sizes):
        psi, n, P, V = data[i]
        p = O.params(N_c, N_r, Lg, Lg, periodic=(True, True), reservoir=flags["reservoir"], pump=flags["pump"],
                     potential=True)
        if flags["reservoir"]:
            o_psi, o_n = O.rk4_run(p, psi, n, pump=P, pot=V, dt=dt, steps=steps)
            assert rel_l2(e.get(i, "n_plus"), o_n) <= 1e-12
        else:
            o_psi, _ = O.rk4_run(p, psi, None, pot=V, dt=dt, steps=steps)
        err = rel_l2(e.get(i, "psi_plus"), o_psi)
        say(f"  instance {i}: rel L2 psi {err:.2e}")
        assert err <= 1e-12
    e.close()
