"""C ABI and host-side logic, CPU only (no device calls).

- libphx.so loads and exports every entry point declared in include/phx.h;
- phx_create parses PHOENIX flags like the reference (aliases, first occurrence
  wins, --config files, FP32 rounding of numbers, derived constants, validation
  errors as status codes instead of exit(1), unknown iterators rejected).
"""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "phx.h")).read()
    return sorted(set(re.findall(r"\b(phx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from phoenix_amd import _lib
    L = _lib.lib()
    syms = header_symbols()
    assert syms, "no declarations found"
    for s in syms:
        assert hasattr(L, s), f"libphx.so does not export {s}"
    assert sorted(_lib.EXPORTED) == syms
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    for s in syms:
        assert re.search(rf"\bT {s}\b", out), s


def make(args, prec="fp64"):
    from phoenix_amd._lib import Solver
    return Solver([str(a) for a in args], prec, initialize=False)


def test_defaults_and_derived_constants():
    s = make(["--N", 100, 60, "--L", 99, 59])
    i = s.info
    assert (i["N_c"], i["N_r"]) == (100, 60)
    assert i["dx"] == pytest.approx(1.0) and i["dy"] == pytest.approx(1.0)  # dx = L/(N-1)
    assert i["dt"] == pytest.approx(0.5 * 1.0 * 1.0)  # magic time step 0.5 dx dy
    assert i["iterator"] == "RK4" and i["halo"] == 4
    assert not i["use_reservoir"]  # no pump, no initial reservoir
    assert i["t_max"] == 1000 and i["fft_every"] == 1


def test_aliases_and_first_occurrence_wins():
    s = make(["--gridsize", 64, 32, "--N", 16, 16, "--tstep", 0.02, "--dt", 0.5, "--boundary", "periodic", "zero"])
    i = s.info
    assert (i["N_c"], i["N_r"]) == (64, 32)
    assert i["dt"] == 0.02
    assert i["periodic_x"] == 1 and i["periodic_y"] == 0


def test_fp32_rounds_numbers():
    s = make(["--N", 32, 32, "--tstep", 0.1, "--tmax", 10], prec="fp32")
    assert s.info["dt"] == float(np.float32(0.1))
    d = make(["--N", 32, 32, "--tstep", 0.1, "--tmax", 10], prec="fp64")
    assert d.info["dt"] == 0.1


def test_config_file_tokens_append_and_cli_wins(tmp_path):
    cfg = tmp_path / "run.cfg"
    cfg.write_text("# comment line\n--N 48 48 # trailing comment\n--tmax 5\n--tstep 0.01\n")
    s = make(["--config", str(cfg), "--tmax", 2])
    i = s.info
    assert (i["N_c"], i["N_r"]) == (48, 48)
    assert i["t_max"] == 2  # command line comes first, first occurrence wins
    assert i["dt"] == 0.01


def test_envelope_flags_enable_reservoir_and_pump():
    s = make(["--N", 32, 32, "--pump", 5.0, "add", 10, 10, 0, 0, "both", 1, "none", "gauss+noDivide"])
    i = s.info
    assert i["use_reservoir"] and i["use_pump"]
    s = make(["--N", 32, 32, "--pump", 5.0, "add", "width", 10, 10, "pos", 0, 0, "-noReservoir"])
    assert not s.info["use_reservoir"]


def test_complex_reservoir_configurations():
    """n is complex under a complex temporal pump amplitude (every iterator) or --imagTime with a reservoir
    (Runge-Kutta iterators); a real pump envelope, imaginary time without a reservoir, or imaginary time on the
    split-step iterator keep it real (the path that serves it is chosen at phx_initialize, on the device)."""
    pump = ["--pump", 1, "add", 10, 10, 0, 0, "both", 1, "none", "gauss"]
    cases = [(pump + ["time", "iexp", 0, 1, 1], 1), (["--imagTime", 1.0] + pump, 1), (pump, 0),
             (pump + ["time", "cos", 0, 1, 1], 0), (["--imagTime", 1.0], 0),
             (["-ssfm", "--imagTime", 1.0] + pump, 0),  # the split-step step takes no -i dt: n stays real
             (["-ssfm"] + pump + ["time", "iexp", 0, 1, 1], 1)]
    for args, want in cases:
        i = make(["--N", 32, 32] + args).info
        assert i["complex_reservoir"] == want, args
        if want:
            assert i["use_reservoir"] == 1


@pytest.mark.parametrize("args,code", [
    (["--N", 31, 32], -1),               # odd dimension
    (["--N", 32, 32, "--tstep", 0], -1),  # dt <= 0
    (["--N", 32, 32, "--tstep", 2, "--tmax", 1], -1),  # dt > t_max
    (["--iterator", "NoSuchKey"], -3),    # unknown key (solver_factory.hpp:59)
    (["--N", "abc", 32], -1),
    (["--pump", 1, "add", 10, 10, 0, 0, "both", 1, "none", "notatype"], -1),
])
def test_invalid_input_returns_status(args, code):
    from phoenix_amd._lib import PhxError
    with pytest.raises(PhxError) as e:
        make(args)
    assert e.value.code == code


def test_adaptive_flag_only_with_adaptive_iterator():
    assert make(["--N", 32, 32, "-adaptive"]).info["adaptive"] == 0  # RK4 cannot adapt (solver.cu:23-26)
    i = make(["--N", 32, 32, "--iterator", "DP45", "-adaptive", "--tol", 1e-6, "--rkvdt", 1e-4, 0.1]).info
    assert i["adaptive"] == 1 and i["halo"] == 7 and i["tolerance"] == 1e-6 and i["dt_max"] == 0.1


# REGISTER_SOLVER keys: (stages fixed, stages adaptive or None)
ITERATORS = {"Newton": (1, None), "MP": (2, None), "Heun2": (2, None), "Ralston": (2, None), "Heun3": (3, None),
             "Ralston3": (3, None), "RK3": (3, None), "SSPRK3": (3, None), "HW": (3, None), "RK4": (4, None),
             "Ralston4": (4, None), "Rule38": (4, None), "Nystroem": (6, None), "Fehlberg2": (3, 3), "BoSha": (3, 4),
             "CashKarp": (6, 6), "Fehlberg5": (6, 6), "DP45": (6, 7),
             # more stages than the fused kernel holds: staged path
             "DP85": (12, 12), "NSRK78": (13, 13)}


@pytest.mark.parametrize("key", list(ITERATORS))
def test_every_registered_iterator_key_is_accepted(key):
    fixed, adaptive = ITERATORS[key]
    i = make(["--N", 32, 32, "--iterator", key.lower()]).info  # keys match without regard to case
    assert i["iterator"] == key and i["halo"] == fixed and i["adaptive"] == 0
    i = make(["--N", 32, 32, "--iterator", key, "-adaptive"]).info
    # iterators registered without variable time stepping run fixed (solver.cu:23-26)
    assert i["adaptive"] == (adaptive is not None) and i["halo"] == (adaptive or fixed)


def test_physics_defaults_match_reference_constants():
    from phoenix_amd._lib import physics_defaults
    ph, magic = physics_defaults(256, 256, 127.5, 127.5)
    hbar_s = 1.0545718e-34 / 1.60217662e-19 * 1e12
    m_eff = 9.10938356e-31 / 1.60217662e-19 * 1e8
    assert ph.one_over_h_bar_s == pytest.approx(1 / hbar_s, rel=1e-15)
    assert ph.m_eff_scaled == pytest.approx(-0.5 * hbar_s ** 2 / m_eff, rel=1e-15)
    assert ph.dx == 0.5 and magic == pytest.approx(0.125)
    # gamma_r = 1.5 * gamma_c and g_r = 2 * g_c as written in system.cpp:29,31
    assert (ph.gamma_c, ph.gamma_r, ph.g_c, ph.g_r, ph.R) == (0.15, 1.5 * 0.15, 3e-6, 2 * 3e-6, 0.01)


def test_ensemble_rejects_invalid_arguments_before_touching_the_device():
    import ctypes as C

    from phoenix_amd import _lib as L
    ph, _ = L.physics_defaults(64, 64, 31.5, 31.5)
    lib = L.lib()
    h = C.c_void_p()

    def create(prec, sizes, flags=3):
        n = len(sizes)
        Nc = (C.c_int * max(n, 1))(*[a for a, _ in sizes])
        Nr = (C.c_int * max(n, 1))(*[b for _, b in sizes])
        return lib.phx_ensemble_create(C.byref(h), prec, C.byref(ph), n, Nc, Nr, None, None, 1, 1, flags)

    assert create(L.PHX_FP32, []) == L.PHX_ERR_INVALID
    assert create(7, [(64, 64)]) == L.PHX_ERR_INVALID
    assert create(L.PHX_FP32, [(64, 64), (0, 12)]) == L.PHX_ERR_INVALID
    assert "instance 1" in lib.phx_last_error().decode()
    assert lib.phx_ensemble_step(None, 0.04, 1, 0) == L.PHX_ERR_INVALID
    assert lib.phx_ensemble_set_tile_rows(None, 12) == L.PHX_ERR_INVALID


def test_executable_rejects_invalid_input_with_exit_code_1(tmp_path):
    """Invalid input exits with status 1 before any device work (system_validate_input.cpp:53-56)."""
    exe = os.path.join(ROOT, "phoenix_amd", "bin", "PHOENIX_gpu_fp32")
    if not os.path.exists(exe):
        pytest.skip("executables not built")
    r = subprocess.run([exe, "--N", "64", "64", "--tstep", "-1", "--path", str(tmp_path)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1
    r = subprocess.run([exe, "--N", "64", "64", "--iterator", "DP87", "--path", str(tmp_path)], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1
    assert "DP87" in (r.stdout + r.stderr)


def test_pyphoenix_config_dict_tokenises_like_the_reference(tmp_path):
    """pyphoenix builds one '--key value ...' string and splits it on whitespace (pyphoenix.py:167-186):
    a key with a trailing space repeats a flag, a value holding a space is two tokens.  The example-3 style
    config of the reference's examples then parses (TE-TM, two initial-state envelopes, a ring pump)."""
    from phoenix_amd.pyphoenix import pyphoenix
    p = pyphoenix(precision="fp32", tetm=True, result_path=str(tmp_path), doprint=False)
    config = {"N": [64, 64], "L": [100, 100], "tmax": 5, "outEvery": 5, "fftEvery": 5, "deltaLT": 0.000025,
              "boundary": ["zero", "zero"], "initRandom": [1E-7, 17],
              "initialState": [1, "add", 3, 3, 0, 0, "plus", 1, 0, "gauss+noDivide+ring"],
              "initialState ": [1, "add", 3, 3, 0, 0, "minus", 1, -4, "gauss+noDivide+ring"],
              "pump": [100, "add", 3, 3, 0, 0, "both", 1, "none gauss+noDivide+ring"], "threads": 2}
    args = p._args(config)
    assert args.count("--initialState") == 2 and "--initialState " not in args
    i = args.index("--pump")
    assert args[i + 9:i + 11] == ["none", "gauss+noDivide+ring"]
    assert args[-4:] == ["--path", ".", "-nosfml", "-tetm"]
    s = make(args, prec="fp32")
    info = s.info
    assert info["tetm"] == 1 and info["use_pump"] == 1 and (info["N_c"], info["periodic_x"]) == (64, 0)
