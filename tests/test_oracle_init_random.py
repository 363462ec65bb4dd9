"""Pins of the oracle's --initRandom restatement (solver_initialization.cu:71-79), CPU only.

  - mt19937: the C++ standard's known answer (the 10000th output of a default-seeded
    std::mt19937 is 4123659995, [rand.predef]);
  - the whole mapping (std::mt19937{seed} + std::uniform_real_distribution<float / double>(-a, a),
    real part first, psi+ then psi-) against a program compiled here with g++ from the standard
    library itself — the library the reference's host code is built with;
  - a planted slip (imaginary part drawn first) is caught.
"""
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

CPP = r"""
#include <cstdio>
#include <cstdlib>
#include <random>
template <typename real> void run(unsigned seed, double amp, int n) {
    std::mt19937 gen{seed};
    std::uniform_real_distribution<real> dist{(real)-amp, (real)amp};
    for (int i = 0; i < n; i++) {
        real a = dist(gen);
        real b = dist(gen);
        std::printf("%a %a\n", (double)a, (double)b);
    }
}
int main(int argc, char** argv) {
    const unsigned seed = (unsigned)std::strtoul(argv[1], nullptr, 10);
    const double amp = std::strtod(argv[2], nullptr);
    const int n = std::atoi(argv[3]);
    if (argv[4][0] == 'f') run<float>(seed, amp, n); else run<double>(seed, amp, n);
}
"""


def test_mt19937_known_answer():
    assert int(O.mt19937_words(5489, 10000)[-1]) == 4123659995


@pytest.fixture(scope="module")
def stdlib_program(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("mt")
    src = d / "u.cpp"
    src.write_text(CPP)
    exe = d / "u"
    subprocess.run([gxx, "-O2", "-std=c++20", str(src), "-o", str(exe)], check=True)
    return exe


def _stdlib(exe, seed, amp, n, kind):
    out = subprocess.run([str(exe), str(seed), repr(amp), str(n), kind], check=True, capture_output=True,
                         text=True).stdout.split()
    v = np.array([float.fromhex(x) for x in out])
    return v[0::2] + 1j * v[1::2]


@pytest.mark.parametrize("fp32", [True, False])
@pytest.mark.parametrize("seed,amp", [(42, 0.01), (7, 1.0), (4294967295, 0.25)])
def test_init_random_equals_stdlib(stdlib_program, fp32, seed, amp):
    Nc, Nr = 9, 7
    ref = _stdlib(stdlib_program, seed, amp, 2 * Nc * Nr, "f" if fp32 else "d")
    plus, minus = O.init_random(Nc, Nr, amp, seed, fp32=fp32, components=2)
    assert np.array_equal(plus.ravel(), ref[:Nc * Nr])
    assert np.array_equal(minus.ravel(), ref[Nc * Nr:])
    # planted slip: imaginary part drawn first
    slip = plus.imag + 1j * plus.real
    assert not np.array_equal(slip.ravel(), ref[:Nc * Nr])
