"""The device index arithmetic, built for the host: closed-form periodic
neighbour steps, multiply-high division by the lattice's invariant divisors,
the tiled work order and the AoSoA slot of a link (its 32-bit offset inside a
class up to the largest lattice), each against plain integer arithmetic over
whole lattices (extents 1 and 2 included) and against x / d for 2e6
numerators."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_index_math(tmp_path):
    exe = tmp_path / "index_math_check"
    subprocess.run([HIPCC, "-std=c++17", "-O1", "-I", os.path.join(ROOT, "ecmc-qncb_amd", "csrc"),
                    "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "index_math_check.cpp"),
                    "-o", str(exe)], check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0/" in r.stdout
