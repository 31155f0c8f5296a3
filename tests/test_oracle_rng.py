"""Pins the oracle's RNG plumbing.

* mt19937_64 + generate_canonical + uniform_int_distribution + normal_distribution
  against the C++ standard library of this image (g++ / libstdc++, the library
  the reference's GCC build uses), via a compiled probe.
* Philox4x32-10 against the Random123 known-answer vectors.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle_lib as qo

PROBE = r"""
#include <random>
#include <cstdio>
int main() {
    std::mt19937_64 g(12345);
    std::uniform_real_distribution<double> u(0, 1);
    for (int i = 0; i < 2000; ++i) std::printf("%.17g\n", u(g));
    std::uniform_int_distribution<int> ui(0, 7);
    for (int i = 0; i < 500; ++i) std::printf("%d\n", ui(g));
    std::uniform_int_distribution<int> ui2(0, 31);
    for (int i = 0; i < 500; ++i) std::printf("%d\n", ui2(g));
    for (int i = 0; i < 300; ++i) {
        std::normal_distribution<double> n(0.0, 1.0);   // fresh object per pair, as random_su3 does
        double a = n(g), b = n(g);
        std::printf("%.17g %.17g\n", a, b);
    }
    return 0;
}
"""


@pytest.fixture(scope="module")
def stdlib_stream(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("probe")
    src = d / "probe.cpp"
    src.write_text(PROBE)
    exe = d / "probe"
    subprocess.run(["g++", "-O2", "-o", str(exe), str(src)], check=True)
    return subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")


def test_mt19937_64_published_value():
    m = qo.Mt64(5489)
    for _ in range(9999):
        m.next()
    assert m.next() == 9981545732273789042


def test_stdlib_distributions(stdlib_stream):
    lines = stdlib_stream
    m = qo.Mt64(12345)
    for i in range(2000):
        assert m.canonical() == float(lines[i])
    for i in range(500):
        assert m.uniform_int(0, 7) == int(lines[2000 + i])
    for i in range(500):
        assert m.uniform_int(0, 31) == int(lines[2500 + i])
    for i in range(300):
        a, b = (float(x) for x in lines[3000 + i].split())
        assert m.normal_pair() == (a, b)


@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_kat(ctr, key, expect):
    assert qo.philox(ctr, key) == expect


def test_u53_open_interval():
    lib = qo.lib()
    assert lib.qo_u53(0, 0) == 2.0 ** -53
    top = lib.qo_u53(0xFFFFFFFF, 0xFFFFFFFF)
    assert 0.0 < top < 1.0 and top == 1.0 - 2.0 ** -53
    assert lib.qo_u53(0x80000000, 0) == 0.5 + 2.0 ** -53
