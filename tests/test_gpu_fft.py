"""GPU parity of the k-space filter (phx_fft.hip) and of the host-evaluated envelopes.

Bars: FP64 max abs error <= 1e-12 x max|psi|, FP32 relative L2 <= 1e-5 against the
FP64 oracle (whose FFT is pinned to numpy.fft in tests/test_oracle.py).
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

MASK_ARGS = ["--fftMask", "1.0", "add", "0.565685", "0.565685", "0", "0", "both", "6", "none", "gauss+noDivide+local"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def make_solver(prec, Nc, Nr, steps=10, dt=0.04, fft_every=None, extra=()):
    from phoenix_amd._lib import Solver
    L = 0.5 * (Nc - 1)
    args = ["--N", Nc, Nr, "--L", L, 0.5 * (Nr - 1), "--boundary", "periodic", "periodic", "--tstep", dt,
            "--tmax", (steps - 0.5) * dt, "--outEvery", 1e9, "--fftEvery", 1e9 if fft_every is None else fft_every]
    args += ["--pump", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide"] + MASK_ARGS + list(extra)
    return Solver([str(a) for a in args], prec)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("shape", [(64, 64), (128, 32), (16, 256), (1024, 1024), (2048, 512), (4096, 32), (8, 8192),
                                   (8, 4096)])
def test_fft_filter_matches_oracle(prec, shape):
    """Row lengths 8 .. 4096 and column lengths 32 .. 8192 cover every column-block width (8, 4, 2, 1)
    and the final radix-2 pass (odd log2 N)."""
    Nc, Nr = shape
    if prec == "fp64" and max(shape) > 4096:
        pytest.skip("FP64 FFT limited to 4096 points")
    say(f"[fft filter] prec={prec} shape={shape}")
    rng = np.random.default_rng(Nc * 7 + Nr)
    psi = rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc))
    mask = rng.random((Nr, Nc))
    s = make_solver(prec, Nc, Nr)
    s.set("psi_plus", psi)
    s.set("fft_mask_plus", mask)
    s.apply_fft_filter()
    g = s.get("psi_plus").astype(np.complex128)
    s.close()
    ref = O.fft_filter(psi, mask)
    if prec == "fp64":
        assert np.abs(g - ref).max() <= 1e-12 * np.abs(psi).max()
    else:
        e = rel_l2(g, ref)
        say(f"  rel L2 = {e:.3e}")
        assert e <= 1e-5


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("shape", [(48, 40), (400, 400), (100, 36), (30, 70), (256, 90), (2048, 6)])
def test_fft_filter_general_sizes_match_oracle(prec, shape):
    """Even sizes that are not powers of two (the reference's default grid is 400 x 400) take the
    direct-DFT kernels; a power of two on one axis only as well."""
    Nc, Nr = shape
    say(f"[fft filter, direct DFT] prec={prec} shape={shape}")
    rng = np.random.default_rng(Nc * 5 + Nr)
    psi = rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc))
    mask = rng.random((Nr, Nc))
    s = make_solver(prec, Nc, Nr)
    s.set("psi_plus", psi)
    s.set("fft_mask_plus", mask)
    s.apply_fft_filter()
    g = s.get("psi_plus").astype(np.complex128)
    s.close()
    ref = O.fft_filter(psi, mask)
    if prec == "fp64":
        e = np.abs(g - ref).max() / np.abs(psi).max()
        say(f"  max abs / max|psi| = {e:.3e}")
        assert e <= 1e-12
    else:
        e = rel_l2(g, ref)
        say(f"  rel L2 = {e:.3e}")
        assert e <= 1e-5


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_fft_filter_analytic(prec):
    """mask = 1 is the identity; a mask passing only k = 0 maps psi to its mean."""
    N = 256
    rng = np.random.default_rng(3)
    psi = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    s = make_solver(prec, N, N)
    s.set("psi_plus", psi)
    s.set("fft_mask_plus", np.ones((N, N)))
    s.apply_fft_filter()
    tol = 1e-12 if prec == "fp64" else 2e-6
    assert np.abs(s.get("psi_plus") - psi).max() <= tol * np.abs(psi).max() * 10
    m = np.zeros((N, N))
    m[N // 2, N // 2] = 1.0  # unshifted mask: k = 0 sits at the centre
    s.set("psi_plus", psi)
    s.set("fft_mask_plus", m)
    s.apply_fft_filter()
    assert np.abs(s.get("psi_plus") - psi.mean()).max() <= tol * 10
    s.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_mask_and_pump_envelopes_match_oracle(prec):
    """Host envelope evaluation (envelope.cpp:332-412) vs the oracle's restatement."""
    N = 64
    L = 0.5 * (N - 1)
    s = make_solver(prec, N, N)
    mask = s.get("fft_mask_plus").astype(np.float64)
    pump = s.get("pump_plus").astype(np.float64)
    s.close()
    dx = L / (N - 1)
    om = O.envelope(N, N, L, L, dx, dx, 1.0, 0.565685, 0.565685, 0, 0, 6.0,
                    O.ENV_GAUSS | O.ENV_NODIVIDE | O.ENV_LOCAL).real
    op = O.envelope(N, N, L, L, dx, dx, 1.0, 10, 10, 0, 0, 1.0, O.ENV_GAUSS | O.ENV_NODIVIDE).real
    # FP32 builds round every command-line number to float before evaluating (commandline_io.cpp:137)
    tol = 1e-14 if prec == "fp64" else 1e-6
    assert np.abs(mask - om).max() <= tol and np.abs(pump - op).max() <= tol


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_rk4_with_periodic_filter_matches_oracle(prec):
    """200 RK4 steps with the filter every 50 steps (fftEvery = 49.5 dt), as in the 1024^2 workload."""
    N, steps, dt = 64, 200, 0.04
    L = 0.5 * (N - 1)
    say(f"[rk4 + filter] prec={prec} N={N}")
    s = make_solver(prec, N, N, steps=steps, dt=dt, fft_every=49.5 * dt)
    psi = O.gaussian_field(N, N, 0.01, 4242)
    x = -L / 2 + L / (N - 1) * np.arange(N)
    X, Y = np.meshgrid(x, x)
    r = np.sqrt(X ** 2 + Y ** 2)
    P = 1.5 * 3.375 * np.exp(-((r - 8.0) ** 2) / (2 * 2.0 ** 2))
    s.set("psi_plus", psi)
    s.set("n_plus", np.zeros((N, N)))
    s.set("pump_plus", P)
    mask = s.get("fft_mask_plus").astype(np.float64)
    assert s.step(steps + 5) == steps
    g = s.get("psi_plus").astype(np.complex128)
    gn = s.get("n_plus").astype(np.float64)
    s.close()
    p = O.params(N, N, L, L, pump=True)
    o, on = psi, np.zeros((N, N))
    for _ in range(steps // 50):
        o, on = O.rk4_run(p, o, on, pump=P, dt=dt, steps=50)
        o = O.fft_filter(o, mask)
    tol = 1e-12 if prec == "fp64" else 1e-5
    e, en = rel_l2(g, o), rel_l2(gn, on)
    say(f"  rel L2 psi {e:.3e} n {en:.3e}")
    assert e <= tol and en <= tol


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("shape,tetm", [((64, 64), False), ((60, 48), True)])
def test_filtered_spectrum_is_kept_for_output(prec, shape, tetm):
    """After a filter application the solver keeps DFT(psi) / (N_c N_r) * fftshift(mask), as the
    reference's fft_plus / fft_minus hold it (solver_fft.cu:67-100); the 60x48 grid takes the direct-DFT
    kernels.  Reference: numpy.fft.fft2 of the same psi."""
    Nc, Nr = shape
    say(f"[fft spectrum, mask] prec={prec} shape={shape} tetm={tetm}")
    rng = np.random.default_rng(Nc + 3 * Nr)
    comps = ["plus", "minus"] if tetm else ["plus"]
    psi = {k: rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc)) for k in comps}
    mask = {k: rng.random((Nr, Nc)) for k in comps}
    s = make_solver(prec, Nc, Nr, steps=10, fft_every=0.3, extra=["-tetm"] if tetm else [])
    for k in comps:
        s.set(f"psi_{k}", psi[k])
        s.set(f"fft_mask_{k}", mask[k])
    s.apply_fft_filter()
    got = {k: s.get(f"fft_{k}").astype(np.complex128) for k in comps}
    s.close()
    for k in comps:
        ref = np.fft.fft2(psi[k]) / (Nc * Nr) * np.fft.fftshift(mask[k])
        e = rel_l2(got[k], ref)
        say(f"  {k}: rel L2 {e:.3e}")
        assert e <= (1e-12 if prec == "fp64" else 1e-5)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_unfiltered_spectrum_of_the_last_fft_time(prec):
    """Without a mask the reference transforms psi at every FFT time for output only (solver_fft.cu:78-79):
    with --fftEvery 1.5 dt the spectrum is DFT(psi) after step 2 (t - t_fft >= fftEvery), not after the
    final step 3, and psi itself is not changed."""
    from phoenix_amd._lib import Solver
    Nc, Nr, dt = 64, 32, 0.04
    say(f"[fft spectrum, no mask] prec={prec}")
    rng = np.random.default_rng(5)
    psi0 = 0.1 * (rng.standard_normal((Nr, Nc)) + 1j * rng.standard_normal((Nr, Nc)))

    def run(steps, fft_every):
        L = 0.5 * (Nc - 1)
        args = ["--N", Nc, Nr, "--L", L, 0.5 * (Nr - 1), "--boundary", "periodic", "periodic", "--tstep", dt,
                "--tmax", (steps - 0.5) * dt, "--outEvery", 1e9, "--fftEvery", fft_every,
                "--pump", 1.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide"]
        s = Solver([str(a) for a in args], prec)
        s.set("psi_plus", psi0)
        assert s.step(steps + 5) == steps
        return s

    s3 = run(3, 1.5 * dt)
    spec = s3.get("fft_plus").astype(np.complex128)
    psi3 = s3.get("psi_plus").astype(np.complex128)
    s3.close()
    s2 = run(2, 1e9)
    psi2 = s2.get("psi_plus").astype(np.complex128)
    s2.close()
    s3b = run(3, 1e9)  # no FFT times at all: psi must be the same
    psi3b = s3b.get("psi_plus").astype(np.complex128)
    s3b.close()
    e = rel_l2(spec, np.fft.fft2(psi2))
    say(f"  rel L2 spectrum {e:.3e}")
    assert e <= (1e-12 if prec == "fp64" else 1e-5)
    assert np.array_equal(psi3, psi3b)
