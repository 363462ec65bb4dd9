"""GPU tests of the command-line and pyphoenix surfaces (the drop-in for the reference's
PHOENIX_gpu_{fp32,fp64} executables and pyphoenix.run).

Bars: the executable's wavefunction_plus.txt / reservoir_plus.txt match the FP64 oracle
(FP64: rel L2 <= 1e-12, FP32: <= 1e-5) for an initial state loaded from a PHOENIX txt
matrix and an envelope pump given on the command line; the summary line is printed;
pyphoenix.run returns 0 and leaves the same files.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "phoenix_amd", "bin")
TOL = {"fp64": 1e-12, "fp32": 1e-5}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def write_txt(path, z, L):
    """PHOENIX txt matrix (filehandler.cpp:282-307 layout): header, real rows, imaginary rows."""
    N_r, N_c = z.shape
    with open(path, "w") as f:
        f.write(f"# SIZE {N_c} {N_r} LX {L} LY {L} DX {L / (N_c - 1)} DY {L / (N_r - 1)} TIME 0 :: PHOENIX_ MATRIX\n")
        for part in (z.real, z.imag):
            for row in part:
                f.write(" ".join(f"{v:.17g}" for v in row) + "\n")


def read_txt(path):
    with open(path) as f:
        head = f.readline().split()
        N_c, N_r = int(head[2]), int(head[3])
        vals = np.array([float(x) for line in f for x in line.split()])
    if vals.size == N_c * N_r:  # a real matrix: one block
        return vals.reshape(N_r, N_c).astype(complex)
    return (vals[:N_c * N_r] + 1j * vals[N_c * N_r:2 * N_c * N_r]).reshape(N_r, N_c)


CASE = dict(N=64, L=31.5, steps=40, dt=0.04)


def cli_args(tmp, prec_path):
    N, L, steps, dt = CASE["N"], CASE["L"], CASE["steps"], CASE["dt"]
    return ["--N", N, N, "--L", L, L, "--boundary", "periodic", "zero", "--tstep", dt, "--tmax", (steps - 0.5) * dt,
            "--outEvery", 1e9, "--fftEvery", 1e9, "--pump", 5.0, "add", 8, 8, 0, 0, "both", 1, "none", "gauss+noDivide",
            "--initialState", "load", prec_path, 1, "add", "both", "--path", str(tmp)]


def oracle_result(psi0):
    N, L, steps, dt = CASE["N"], CASE["L"], CASE["steps"], CASE["dt"]
    dx = L / (N - 1)
    P = O.envelope(N, N, L, L, dx, dx, 5.0, 8, 8, 0, 0, 1.0, 1 | 8).real
    p = O.params(N, N, L, L, periodic=(True, False), reservoir=True, pump=True)
    return O.rk4_run(p, psi0, np.zeros((N, N)), pump=P, dt=dt, steps=steps)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_executable_matches_oracle(prec, tmp_path):
    N, L = CASE["N"], CASE["L"]
    say(f"[cli] PHOENIX_gpu_{prec} {N}x{N}")
    psi0 = O.gaussian_field(N, N, 0.2, 31)
    init = tmp_path / "psi0.txt"
    write_txt(init, psi0, L)
    exe = os.path.join(BIN, f"PHOENIX_gpu_{prec}")
    r = subprocess.run([exe] + [str(a) for a in cli_args(tmp_path / "out", init)], capture_output=True, text=True,
                       timeout=120)
    say(r.stdout[-400:])
    assert r.returncode == 0, r.stderr
    assert "Total Runtime" in r.stdout
    g_psi = read_txt(tmp_path / "out" / "wavefunction_plus.txt")
    g_n = read_txt(tmp_path / "out" / "reservoir_plus.txt").real
    o_psi, o_n = oracle_result(psi0)
    e_psi, e_n = rel_l2(g_psi, o_psi), rel_l2(g_n, o_n)
    say(f"  rel L2 psi {e_psi:.3e} n {e_n:.3e}")
    assert e_psi <= TOL[prec] and e_n <= TOL[prec]


def test_pyphoenix_run_matches_executable(tmp_path):
    from phoenix_amd.pyphoenix import pyphoenix
    N, L = CASE["N"], CASE["L"]
    say("[pyphoenix] run fp32")
    psi0 = O.gaussian_field(N, N, 0.2, 32)
    init = tmp_path / "psi0.txt"
    write_txt(init, psi0, L)
    args = cli_args(tmp_path / "unused", init)
    config = {}
    i = 0
    while i < len(args):  # flag -> values, as pyphoenix configs are written
        key = args[i][2:]
        j = i + 1
        while j < len(args) and not str(args[j]).startswith("--"):
            j += 1
        if key != "path":
            vals = [str(v) for v in args[i + 1:j]]
            config[key] = vals if len(vals) > 1 else vals[0]
        i = j
    p = pyphoenix(precision="fp32", result_path=str(tmp_path / "py"), doprint=False)
    assert p.have_bin
    assert p.run(config) == 0, p.stderr
    g_psi = read_txt(tmp_path / "py" / "wavefunction_plus.txt")
    o_psi, _ = oracle_result(psi0)
    e = rel_l2(g_psi, o_psi)
    say(f"  rel L2 psi {e:.3e}")
    assert e <= TOL["fp32"]


def test_executable_writes_the_reference_outputs(tmp_path):
    """Default --output keys ("mat,scalar"): initial matrices (one file per temporal group of a
    source, with its oscillator header), the final matrices and scalar.txt with one row per
    output interval; --historyMatrix snapshots under timeoutput/ (solver_cache.cu:7-72,
    solver_output_matrices.cu:20-150)."""
    N, L, dt = 32, 15.5, 0.04
    exe = os.path.join(BIN, "PHOENIX_gpu_fp64")
    out = tmp_path / "out"
    t0, sigma, freq = 0.5, 0.3, 4.0
    args = ["--N", N, N, "--L", L, L, "--boundary", "periodic", "periodic", "--tstep", dt, "--tmax", 1.0,
            "--outEvery", 0.2, "--fftEvery", 1e9, "--initRandom", 0.1, 5,
            "--pump", 5.0, "add", 8, 8, 0, 0, "both", 1, "none", "gauss+noDivide",
            "--pump", 2.0, "add", 4, 4, 3, 0, "both", 1, "none", "gauss+noDivide", "time", "cos", t0, sigma, freq,
            "--historyMatrix", 0, 16, 0, 16, 2, "--path", str(out)]
    say("[cli outputs] PHOENIX_gpu_fp64 32x32, outEvery 0.2, two pump groups, history matrix")
    r = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for line in ("T = ", "Progress:", "Current System: Scalar", "Runtime:", "Time per ps:", "Total Runtime"):
        assert line in r.stdout, line
    names = sorted(os.listdir(out))
    say("  files: " + " ".join(names))
    for f in ("initial_wavefunction_plus.txt", "initial_reservoir_plus.txt", "pump_plus.txt", "pump_plus_1.txt",
              "wavefunction_plus.txt", "reservoir_plus.txt", "scalar.txt", "timeoutput"):
        assert f in names, f
    dx = L / (N - 1)
    P0 = O.envelope(N, N, L, L, dx, dx, 5.0, 8, 8, 0, 0, 1.0, 1 | 8)
    P1 = O.envelope(N, N, L, L, dx, dx, 2.0, 4, 4, 3, 0, 1.0, 1 | 8)
    assert np.abs(read_txt(out / "pump_plus.txt") - P0).max() < 1e-12 * np.abs(P0).max()
    assert np.abs(read_txt(out / "pump_plus_1.txt") - P1).max() < 1e-12 * np.abs(P1).max()
    head = open(out / "pump_plus_1.txt").readline()
    assert f"OSC T0 {t0:g} FREQ {freq:g} SIGMA {sigma:g}" in head, head
    # scalar.txt: keys in std::map order, one row per output interval
    lines = open(out / "scalar.txt").read().split("\n")
    cols = lines[0].split()
    assert cols == ["index", "max_plus", "min_plus", "pump_1", "t"], cols
    rows = np.array([[float(x) for x in ln.split()] for ln in lines[1:] if ln.strip()])
    assert rows.shape[0] == 5 and np.allclose(rows[:, 4], [0.2, 0.4, 0.6, 0.8, 1.0], atol=1e-12)
    tt = rows[:, 4]
    pump_1 = np.exp(-0.5 * ((tt - t0) / sigma) ** 2) * (1 + np.cos(freq * (tt - t0))) / 2
    assert np.allclose(rows[:, 3], pump_1, rtol=1e-9, atol=1e-12)
    psi = read_txt(out / "wavefunction_plus.txt")
    assert abs(rows[-1, 1] - np.abs(psi).max()) < 1e-9 * np.abs(psi).max()
    assert abs(rows[-1, 2] - np.abs(psi).min()) < 1e-9 * np.abs(psi).max()
    hist = sorted(os.listdir(out / "timeoutput"))
    assert len([h for h in hist if h.startswith("wavefunction_plus_")]) == 5, hist
    h = read_txt(out / "timeoutput" / [x for x in hist if x.startswith("wavefunction_plus_1.0")][0])
    assert h.shape == (8, 8) and np.abs(h - psi[0:16:2, 0:16:2]).max() < 1e-12 * np.abs(psi).max()


def test_executable_name_prefix_and_verbose_io(tmp_path):
    """--name X: every output file is <path>X_<file>, the history directory X_timeoutput/ included
    (filehandler.cpp:40-41,461-463); -verboseIO reports each written matrix (filehandler.cpp:309-312)."""
    exe = os.path.join(BIN, "PHOENIX_gpu_fp64")
    out = tmp_path / "out"
    args = ["--N", 32, 32, "--L", 15.5, 15.5, "--tstep", 0.04, "--tmax", 0.4, "--outEvery", 0.2, "--fftEvery", 1e9,
            "--pump", 5.0, "add", 8, 8, 0, 0, "both", 1, "none", "gauss+noDivide", "--historyMatrix", 0, 8, 0, 8, 1,
            "--outputPath", str(out), "--name", "run7", "-verboseIO"]
    say("[cli --name / -verboseIO] PHOENIX_gpu_fp64 32x32")
    r = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    names = sorted(os.listdir(out))
    say("  files: " + " ".join(names))
    for f in ("run7_initial_wavefunction_plus.txt", "run7_wavefunction_plus.txt", "run7_reservoir_plus.txt",
              "run7_scalar.txt", "run7_timeoutput"):
        assert f in names, f
    assert not any(n.startswith(("wavefunction", "scalar")) for n in names)
    assert f"Output 1024 elements to '{out}/run7_wavefunction_plus.txt'." in r.stdout
    assert os.listdir(out / "run7_timeoutput")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_executable_binary_output_types(prec, tmp_path):
    """--outputFileType npy / ash / mat (filehandler.cpp:49-69, simple_binary.hpp): the same matrices
    as the txt output, in the build's precision."""
    N, L, dt = 32, 15.5, 0.04
    exe = os.path.join(BIN, f"PHOENIX_gpu_{prec}")
    base = ["--N", N, N, "--L", L, L, "--boundary", "periodic", "zero", "--tstep", dt, "--tmax", 0.38,
            "--outEvery", 1e9, "--fftEvery", 1e9, "--initRandom", 0.1, 9,
            "--pump", 5.0, "add", 8, 8, 0, 0, "both", 1, "none", "gauss+noDivide"]
    res = {}
    for ft in ("txt", "npy", "ash", "mat"):
        out = tmp_path / ft
        r = subprocess.run([exe] + [str(a) for a in base + ["--outputFileType", ft, "--path", out]], capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        res[ft] = out
    say(f"[cli binary outputs] {prec}: " + " ".join(sorted(os.listdir(res['npy']))))
    psi = read_txt(res["txt"] / "wavefunction_plus.txt")
    pump = read_txt(res["txt"] / "pump_plus.txt").real
    ctype, rtype = (np.complex64, np.float32) if prec == "fp32" else (np.complex128, np.float64)
    a = np.load(res["npy"] / "wavefunction_plus.npy")
    assert a.dtype == ctype and a.shape == (N, N)
    assert np.array_equal(a.astype(np.complex128), psi) if prec == "fp64" else np.abs(a - psi).max() < 1e-6 * np.abs(psi).max()
    b = np.load(res["npy"] / "pump_plus.npy")
    assert b.dtype == rtype and np.abs(b - pump).max() <= 1e-6 * np.abs(pump).max()
    # ash: the npy layout with the magic "\x93PHOENIX"
    raw = open(res["ash"] / "wavefunction_plus.ash", "rb").read()
    assert raw[:8] == b"\x93PHOENIX" and raw[8:10] == b"\x01\x00"
    hl = raw[10] | (raw[11] << 8)
    data = np.frombuffer(raw[12 + hl:], dtype=ctype)
    assert np.array_equal(data.reshape(N, N), a)
    # mat: 128-byte text, 5 int32 (type, rows, cols, complex, name length), "mat\0", real then imaginary block,
    # column-major
    raw = open(res["mat"] / "wavefunction_plus.mat", "rb").read()
    hdr = np.frombuffer(raw[128:148], dtype=np.int32)
    assert list(hdr) == [0, N, N, 1, 4] and raw[148:152] == b"mat\x00"
    vals = np.frombuffer(raw[152:], dtype=rtype)
    re_, im_ = vals[:N * N].reshape(N, N).T, vals[N * N:].reshape(N, N).T
    assert np.array_equal(re_ + 1j * im_, a)


def test_reference_example_2_config_through_pyphoenix(tmp_path):
    """The pyphoenix config of the reference's example 2 (examples/example_2/example.ipynb: 200 x 200 grid,
    loaded pump / initial state / initial reservoir / potential, a local super-Gaussian k-space mask applied
    every step -- a non-power-of-two grid, so the direct-DFT filter) for a few steps, against the oracle
    stepping RK4 and filtering after every step.  Deviations from the notebook: no random initial noise
    (its seed is random) and a short t_max."""
    from phoenix_amd.pyphoenix import pyphoenix
    N, x_max, m_eff, g_c, gamma_c, R, dt, steps = 200, 20, 0.5 * 5.864E-4, 6E-6, 0.16, 0.01, 0.02, 6
    dx = x_max / (N - 1)
    x = -x_max / 2 + dx * np.arange(N)
    X, Y = np.meshgrid(x, x)
    r = np.sqrt(X ** 2 + Y ** 2)
    P = 20.0 * np.exp(-((r - 4.0) / 1.5) ** 2)
    V = 2e-3 * (np.exp(-((X - 2) ** 2 + Y ** 2) / 4.0) - np.exp(-((X + 2) ** 2 + Y ** 2) / 4.0))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        write_txt(tmp_path / "pump_plus.txt", P.astype(complex), x_max)
        write_txt(tmp_path / "potential_plus.txt", V.astype(complex), x_max)
        config = {"N": [N, N], "L": [x_max, x_max], "meff": m_eff, "tmax": (steps - 0.5) * dt, "tstep": dt,
                  "outEvery": 1e9, "gc": g_c, "gr": 2 * g_c, "R": R, "gammaR": 1.5 * gamma_c, "gammaC": gamma_c,
                  "boundary": ["periodic", "periodic"], "fftEvery": 0.001,
                  "fftMask": [1.0, "add", 0.2, 0.2, 0, 0, "both", 10, "none", "gauss+local+noDivide"],
                  "pump": ["load", "pump_plus.txt", "1", "add", "both"],
                  "initialState": ["load", "pump_plus.txt", 1, "add", "both"],
                  "initialReservoir": ["load", "pump_plus.txt", 1, "add", "both"],
                  "potential": ["load", "potential_plus.txt", 1, "add", "both"]}
        say("[reference example 2] 200x200 fp32 via pyphoenix, k-space mask every step")
        p = pyphoenix(precision="fp32", result_path=str(tmp_path / "out"), doprint=False)
        assert p.run(config) == 0, p.stderr
    finally:
        os.chdir(cwd)
    g = read_txt(tmp_path / "out" / "wavefunction_plus.txt")
    mask = O.envelope(N, N, x_max, x_max, dx, dx, 1.0, 0.2, 0.2, 0, 0, 10.0, 1 | 8 | 16).real
    par = O.params(N, N, x_max, x_max, periodic=(True, True), pump=True, potential=True, m_eff=m_eff, g_c=g_c,
                   g_r=2 * g_c, R=R, gamma_r=1.5 * gamma_c, gamma_c=gamma_c)
    psi, n = P.astype(complex), P.copy()
    for _ in range(steps):
        psi, n = O.rk4_run(par, psi, n, pump=P, pot=V, dt=dt, steps=1)
        psi = O.fft_filter(psi, mask)
    e = rel_l2(g, psi)
    say(f"  rel L2 psi {e:.3e}")
    assert e <= 1e-5


def test_executable_writes_the_kspace_spectrum(tmp_path):
    """With --fftEvery below --tmax the default keys write fft_plus (solver_output_matrices.cu:41): the
    spectrum of the last FFT time, complex, with the k-space header LX = LY = -1, DX = 2/N_c, DY = 2/N_r."""
    N, L, dt = 32, 15.5, 0.04
    exe = os.path.join(BIN, "PHOENIX_gpu_fp64")
    out = tmp_path / "out"
    args = ["--N", N, N, "--L", L, L, "--boundary", "periodic", "periodic", "--tstep", dt, "--tmax", 0.5,
            "--outEvery", 1e9, "--fftEvery", 0.2, "--initRandom", 0.1, 5,
            "--pump", 5.0, "add", 8, 8, 0, 0, "both", 1, "none", "gauss+noDivide",
            "--fftMask", 1.0, "add", 0.5, 0.5, 0, 0, "both", 6, "none", "gauss+noDivide+local", "--path", str(out)]
    say("[cli k-space output] PHOENIX_gpu_fp64 32x32, fftEvery 0.2")
    r = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    head = open(out / "fft_plus.txt").readline()
    assert head.startswith(f"# SIZE {N} {N} LX -1 LY -1 DX {2 / N:g} DY {2 / N:g} TIME "), head
    F = read_txt(out / "fft_plus.txt")
    mask = np.fft.fftshift(read_txt(out / "fft_mask_plus.txt").real)
    assert F.shape == (N, N) and np.isfinite(F).all() and np.abs(F).max() > 0
    # the filtered spectrum vanishes where the (shifted) mask does
    assert np.abs(F[mask < 1e-12]).max(initial=0.0) <= 1e-12 * np.abs(F).max()
