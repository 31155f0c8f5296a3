"""The gauge_hb_cb driver's input parser (reference src/io/io.cpp:353-399
semantics) without a GPU: key=value with trimming, '#' comments at column 0,
unknown keys ignored, cold_start true only for the literal "true",
N_samples forced to 1, N_switch_eo never parsed."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "ecmc-qncb_amd", "gauge_hb_cb")

INPUT = """#Lattice params
L_core=4
n_core_dims=2
cold_start=false

#Run params
seed=12345
N_shift_therm = 0
N_shift=50000

#Heatbath params
beta = 6.0
N_sweeps = 2
N_hits = 1
N_switch_eo = 9
  # indented comment lines are not comments: no '=' so they are skipped
unknown_key = 3

#Plaquette params
N_shift_plaquette = 1

#Topo params
topo = false
N_shift_topo = 25
N_steps_gf = 10
N_rk_steps = 40

#Save params
run_name = hhot62
run_dir = data
save_each_shifts = 2000
"""


def run(tmp_path, text):
    f = tmp_path / "in.txt"
    f.write_text(text)
    r = subprocess.run([DRIVER, "--print-params", str(f)], capture_output=True, text=True, timeout=30)
    return r


@pytest.fixture(autouse=True)
def _need_driver():
    if not os.path.exists(DRIVER):
        pytest.fail("gauge_hb_cb is not built (make -C ecmc-qncb_amd)")


def test_reference_input_template(tmp_path):
    r = run(tmp_path, INPUT)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert "Heatbath - MPI No Checkboard" in out
    assert "Run name : hhot62" in out
    assert "Total lattice size : 8^4" in out and "Local lattice size : 4^4" in out
    assert "Initial seed : 12345" in out
    assert "Thermalization shifts : 0" in out and "Number of shifts : 50000" in out
    assert "Beta : 6\n" in out and "Number of sweeps : 2" in out and "Number of hits : 1" in out
    assert "Measure topo : No" in out
    assert "N_samples=1" in out and "N_switch_eo=3" in out   # forced / never parsed
    assert "precision=fp64 reunit_every=-1" in out


def test_defaults_and_cold_start_literal(tmp_path):
    r = run(tmp_path, "cold_start=hot\n")
    out = r.stdout
    assert r.returncode == 0
    # defaults of RunParamsHbCB (src/io/params.h:38-56)
    assert "Total lattice size : 12^4" in out and "Run name : c" in out
    assert "Measure Q each : 10 shifts" in out and "Number of GF steps : 10" in out


def test_missing_file(tmp_path):
    r = subprocess.run([DRIVER, "--print-params", str(tmp_path / "nope.txt")], capture_output=True, text=True)
    assert r.returncode == 1 and "Error reading input : cannot open input file" in r.stderr


def test_precision_keys(tmp_path):
    """The driver's own keys: precision = fp64 | fp32 (fp32 reunitarises every
    4 sweeps unless reunit_every says otherwise); other values are refused."""
    r = run(tmp_path, "precision = fp32\n")
    assert r.returncode == 0 and "precision=fp32 reunit_every=-1" in r.stdout
    r = run(tmp_path, "precision = fp32\nreunit_every = 10\n")
    assert r.returncode == 0 and "precision=fp32 reunit_every=10" in r.stdout
    for bad in ("precision = fp16\n", "reunit_every = -2\n", "reunit_every = 3x\n", "reunit_every = four\n"):
        r = run(tmp_path, bad)
        assert r.returncode == 1 and "Error reading input" in r.stderr, bad
