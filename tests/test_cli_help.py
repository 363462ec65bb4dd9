"""The executables' --help / -h, --helpVerbose / -hv and --markdown / -mk (system.cpp:162-170: print the
flags and exit 0 before any parsing or device work).  CPU only."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "phoenix_amd", "bin")


@pytest.mark.parametrize("exe", ["PHOENIX_gpu_fp32", "PHOENIX_gpu_fp64"])
def test_help_lists_the_flags_and_exits_zero(exe):
    path = os.path.join(BIN, exe)
    r = subprocess.run([path, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    for flag in ("--N", "--tmax", "--iterator", "--imagTime", "--pump", "--path", "--name", "--outputFileType"):
        assert flag in r.stdout, flag
    assert "-verboseIO" not in r.stdout  # verbose-only entries
    v = subprocess.run([path, "-hv"], capture_output=True, text=True, timeout=60)
    assert v.returncode == 0 and "-verboseIO" in v.stdout and "Envelope syntax" in v.stdout
    m = subprocess.run([path, "-h", "-mk"], capture_output=True, text=True, timeout=60)
    assert m.returncode == 0 and "| Flag | Arguments | Meaning |" in m.stdout and "| `--name` |" in m.stdout
