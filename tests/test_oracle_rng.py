"""Oracle random-number generators against published known answers.

- mt19937_64: the C++11 standard fixes the 10000th output of a
  default-seeded engine ([rand.predef]).
- libstdc++ uniform_real_distribution<double>: compared draw for draw with a
  small program compiled by this image's g++ (its <random>, the library GLAC
  calls at GLAC/system.cpp:40-41, 347 and
  GLAC/math/matrices/su3matrixgenerator.cpp:22-24).
- Philox4x32-10: known-answer vectors published with Random123 (Salmon et
  al., SC'11).
- xoshiro128** 1.1: the first ten outputs from the state (1, 2, 3, 4), as
  produced by Blackman & Vigna's reference C code (the same vector is the
  known-answer test of the Rust rand_xoshiro crate); its output statistics are
  checked as well.
"""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O


def test_mt19937_64_10000th_output():
    g = O.MT64(5489)
    for _ in range(9999):
        g.next()
    assert g.next() == 9981545732273789042


CPP = r"""
#include <cstdio>
#include <random>
int main() {
    std::mt19937_64 g(12345);
    std::uniform_real_distribution<double> a(-0.5, 0.5), b(0.0, 1.0), c(-0.24, 0.24);
    for (int i = 0; i < 300; i++) {
        double x = a(g), y = b(g), z = c(g);
        std::printf("%.17g %.17g %.17g\n", x, y, z);
    }
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_uniform_real_distribution_matches_libstdcxx():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "u.cpp")
        exe = os.path.join(d, "u")
        with open(src, "w") as f:
            f.write(CPP)
        subprocess.run(["g++", "-O1", "-std=c++11", src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    want = np.array([[float(v) for v in line.split()] for line in out.strip().splitlines()])
    g = O.MT64(12345)
    got = np.array([[g.uniform(-0.5, 0.5), g.uniform(0.0, 1.0), g.uniform(-0.24, 0.24)]
                    for _ in range(300)])
    assert np.array_equal(got, want)


@pytest.mark.parametrize("ctr,key,want", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox4x32_10_known_answers(ctr, key, want):
    assert list(O.philox4x32_10(ctr, key)) == want


def test_xoshiro128ss_known_answer():
    st = np.array([1, 2, 3, 4], dtype=np.uint32)
    want = [11520, 0, 5927040, 70819200, 2031721883, 1637235492,
            1287239034, 3734860849, 3729100597, 4258142804]
    assert [int(v) for v in O.xoshiro128ss(st, 10)] == want


def test_xoshiro128ss_output_is_uniform():
    st = O.philox4x32_10([7, 0, 0, 0], [1, 2]).copy()
    x = O.xoshiro128ss(st, 20000).astype(np.float64) / 2.0**32
    counts, _ = np.histogram(x, bins=20, range=(0, 1))
    chi2 = ((counts - 1000.0) ** 2 / 1000.0).sum()
    assert chi2 < 50.0  # 19 dof, p ~ 1e-4
    assert abs(np.corrcoef(x[:-1], x[1:])[0, 1]) < 0.03
