"""CPU pinning of the rotation draws of the Monte-Carlo 3PCF.

The reference draws each rotation's uniforms with Numba's ``np.random.seed(s); np.random.rand(3)``
(corr3pcf_kernels.py:119-125).  The oracle restates Numba's generator (MT19937, init_genrand seeding,
genrand_res53 doubles) in oracle/mt19937.py; it is pinned here by the C++ standard's known value and by
libstdc++'s independent std::mt19937, and the product's draws are checked against it bit for bit.
"""

import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import mc, mt19937

STD_MT_PROGRAM = r"""
#include <cstdio>
#include <cstdlib>
#include <random>
int main(int argc, char** argv) {
    unsigned long seed = std::strtoul(argv[1], nullptr, 10);
    int n = std::atoi(argv[2]);
    std::mt19937 g((std::mt19937::result_type)seed);
    for (int i = 0; i < n; ++i) std::printf("%u\n", (unsigned)g());
    return 0;
}
"""


def test_mt19937_known_values():
    """std::mt19937 seeded 5489 (its default seed): first output 3499211612, 10,000th output 4123659995
    (the value the C++ standard requires, [rand.predef])."""
    g = mt19937.MT19937(5489)
    first = g.genrand_int32()
    for _ in range(9998):
        g.genrand_int32()
    assert first == 3499211612
    assert g.genrand_int32() == 4123659995


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ for libstdc++'s std::mt19937")
def test_mt19937_matches_libstdcxx(tmp_path):
    """1500 outputs (three regenerations of the 624-word state) for seeds at the edges of the 32-bit range
    and at the rotation-seed offsets the task uses, against libstdc++'s std::mt19937."""
    src = tmp_path / "mt.cpp"
    src.write_text(STD_MT_PROGRAM)
    exe = tmp_path / "mt"
    subprocess.run(["g++", "-O1", "-std=c++11", str(src), "-o", str(exe)], check=True)
    for seed in (0, 1, 5489, 8 + 2 * 1000003, 2 ** 32 - 1):
        out = subprocess.run([str(exe), str(seed), "1500"], check=True, capture_output=True, text=True).stdout
        want = [int(v) for v in out.split()]
        g = mt19937.MT19937(seed)
        assert [g.genrand_int32() for _ in range(1500)] == want, f"seed {seed}"


def test_res53_doubles_from_int32_pairs():
    """genrand_res53 = (a >> 5, b >> 6) of two consecutive outputs: (a * 2^26 + b) / 2^53, in [0, 1)."""
    g, h = mt19937.MT19937(77), mt19937.MT19937(77)
    for _ in range(500):
        a, b = h.genrand_int32() >> 5, h.genrand_int32() >> 6
        x = g.genrand_res53()
        assert x == (a * 2.0 ** 26 + b) / 2.0 ** 53
        assert 0.0 <= x < 1.0


def test_product_rotation_draws_match_numba_generator():
    """The task's rotation draws (numpy's legacy RandomState(s).rand(3), seeds of corr3pcf.py:1462) equal the
    restated Numba stream bit for bit, for several angle indices and rotations."""
    from pyhermes_amd.theory.corr3pcf import rotation_draws
    for base, n_rot, it in ((1, 6, 0), (43, 5, 2), (1000, 4, 7), (2 ** 31, 3, 1)):
        np.testing.assert_array_equal(rotation_draws(base, n_rot, it), mc.rotation_uniforms(base, n_rot, it))
