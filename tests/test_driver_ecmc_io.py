"""The gauge_ecmc_cb driver's host side without a GPU: the input parser
(reference src/io/io.cpp:298-349 semantics) and the per-rank chain-state
files it resumes from (src/io/io.cpp:160-232).  State file r belongs to the
rank whose block MPI_Cart_create gives it (row-major, t fastest); its site is
the reference's padded index ((t*Le + z)*Le + y)*Le + x with Le = L_core + 2
and the block's sites at 1..L_core.  Files in the reference's own layout
(set_counter and events written with no separator, no Philox line) load with
both counters and the Philox position zeroed."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "ecmc-qncb_amd", "gauge_ecmc_cb")

TEMPLATE = """#Lattice params
L_core=4
n_core_dims=2
cold_start=hot

#Run params
seed=12345
N_shift_therm = 0
N_shift=50000

#ECMC params
beta=5
param_theta_sample=1800
param_theta_refresh_site=1500
param_theta_refresh_R=150
poisson=false
epsilon_set=0.15

#Plaquette params
N_shift_plaquette = 1

#Topo params
topo = false
N_shift_topo = 25
N_steps_gf = 10
N_rk_steps = 40

#Save params
run_name = ehot62
run_dir = {run_dir}
save_each_shifts = 2000
"""


@pytest.fixture(autouse=True)
def _need_driver():
    if not os.path.exists(DRIVER):
        pytest.fail("gauge_ecmc_cb is not built (make -C ecmc-qncb_amd)")


def run(args):
    return subprocess.run([DRIVER] + args, capture_output=True, text=True, timeout=30)


def test_reference_ecmc_template(tmp_path):
    f = tmp_path / "in.txt"
    f.write_text(TEMPLATE.format(run_dir=tmp_path / "data"))
    r = run(["--print-params", str(f)])
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert "ECMC - MPI No Checkboard" in out and "Run name : ehot62" in out
    assert "Total lattice size : 8^4" in out and "Local lattice size : 4^4" in out
    assert "Initial seed : 12345" in out and "Thermalization shifts : 0" in out
    assert "Number of shifts : 50000" in out and "Save each : 2000 shifts" in out
    assert "Beta : 5\n" in out and "Theta sample : 1800" in out and "Theta refresh site : 1500" in out
    assert "Theta refresh R : 150" in out and "Epsilon set : 0.15" in out and "Poisson law : No" in out
    assert "Measure topo : No" in out and "Number of <P> samples : 50000" in out
    # "hot" is not the literal "true": the reference parser leaves a hot start
    assert "N_samples=1 N_switch_eo=3 cold_start=false" in out
    assert not (tmp_path / "data").exists()   # parsing creates no files


def test_defaults_and_missing_file(tmp_path):
    f = tmp_path / "in.txt"
    f.write_text("# nothing set\n")
    r = run(["--print-params", str(f)])
    out = r.stdout
    # defaults of RunParamsECB / ECMCParams (src/io/params.h:9-44)
    assert "Total lattice size : 12^4" in out and "Run name : c" in out
    assert "Theta sample : 100" in out and "Theta refresh site : 50" in out and "Theta refresh R : 15" in out
    assert "Thermalization shifts : 100" in out and "Number of shifts : 2" in out
    assert "N_samples=1 N_switch_eo=3 cold_start=true" in out
    r = run(["--print-params", str(tmp_path / "nope.txt")])
    assert r.returncode == 1 and "Error reading input : cannot open input file" in r.stderr


def rank_block(r, n):
    return [(r // n ** (3 - d)) % n for d in range(4)]


def write_states(tmp_path, glued_ranks, rng):
    """16 state files of an 8^4 run in 4^4 blocks; returns what each rank holds"""
    Lc, n = 4, 2
    Le = Lc + 2
    d = tmp_path / "data" / "ehot62" / "ehot62_state"
    d.mkdir(parents=True)
    want = {}
    for r in range(n ** 4):
        loc = rng.integers(1, Lc + 1, size=4)
        idx = ((loc[3] * Le + loc[2]) * Le + loc[1]) * Le + loc[0]
        s = dict(mu=int(rng.integers(0, 4)), eps=int(rng.choice([-1, 1])), par_rs=float(rng.uniform(0, 50)),
                 par_rR=float(rng.uniform(0, 15)), th_rs=50.0, th_rR=15.0, set_counter=int(rng.integers(1, 9)),
                 events=int(rng.integers(10, 99)), lifts=int(rng.integers(1, 9)), revs=int(rng.integers(0, 5)),
                 init=1, R=rng.normal(size=18), rng_n=int(rng.integers(100, 1000)), set_gen=int(rng.integers(2, 9)))
        head = [idx, s["mu"], s["eps"], repr(s["par_rs"]), repr(s["par_rR"]), repr(s["th_rs"]), repr(s["th_rR"])]
        if r in glued_ranks:
            # the reference's writer: set_counter and events with no separator
            head += [f"{s['set_counter']}{s['events']}", s["lifts"], s["revs"], "", s["init"]]
            extra = ""
            s.update(set_counter=0, events=0, rng_n=0, set_gen=0)
        else:
            head += [s["set_counter"], s["events"], s["lifts"], s["revs"], s["init"]]
            extra = f"{s['rng_n']} {s['set_gen']}\n"
        rows = "".join(" ".join(repr(float(v)) for v in s["R"][6 * i:6 * i + 6]) + " \n" for i in range(3))
        (d / f"ehot62_state{r}.txt").write_text(" ".join(str(v) for v in head) + "\n" + rows + extra)
        b = rank_block(r, n)
        s["site"] = [b[k] * Lc + int(loc[k]) - 1 for k in range(4)]
        want[r] = s
    return want


def parse_states(out):
    got = {}
    for line in out.strip().splitlines():
        t = line.split()
        if not t or t[0] != "chain":   # the loader's own "Successfully loaded ..." line
            continue
        c = int(t[1])
        site = [int(x) for x in t[3:7]]
        kv = {t[i]: t[i + 1] for i in range(7, 7 + 2 * 13, 2)}
        R = np.array([float(x) for x in t[t.index("R") + 1:]])
        got[c] = dict(site=site, R=R, **{k: float(x) for k, x in kv.items()})
    return got


@pytest.mark.parametrize("glued", [(), (0, 3, 5, 10, 15)])
def test_state_files_by_rank(tmp_path, glued):
    rng = np.random.default_rng(7 + len(glued))
    want = write_states(tmp_path, set(glued), rng)
    f = tmp_path / "in.txt"
    f.write_text(TEMPLATE.format(run_dir=tmp_path / "data"))
    r = run(["--print-states", str(f)])
    assert r.returncode == 0, r.stderr
    got = parse_states(r.stdout)
    assert len(got) == 16
    n = 2
    for rank, s in want.items():
        b = rank_block(rank, n)
        c = b[0] + n * (b[1] + n * (b[2] + n * b[3]))   # device chains: x fastest
        g = got[c]
        assert g["site"] == s["site"]
        assert all(s["site"][k] // 4 == b[k] for k in range(4))   # the rank's own block
        assert (g["mu"], g["eps"], g["init"]) == (s["mu"], s["eps"], s["init"])
        assert (g["set_counter"], g["events"], g["lifts"], g["revs"]) == (s["set_counter"], s["events"], s["lifts"],
                                                                          s["revs"])
        assert (g["rng_n"], g["set_gen"]) == (s["rng_n"], s["set_gen"])
        for k, kk in (("par_rs", "par_rs"), ("par_rR", "par_rR"), ("th_rs", "th_rs"), ("th_rR", "th_rR")):
            assert g[kk] == s[k]
        assert np.array_equal(g["R"], s["R"])


def test_state_files_rejected(tmp_path):
    rng = np.random.default_rng(3)
    write_states(tmp_path, set(), rng)
    d = tmp_path / "data" / "ehot62" / "ehot62_state"
    f = tmp_path / "in.txt"
    f.write_text(TEMPLATE.format(run_dir=tmp_path / "data"))
    p = d / "ehot62_state4.txt"
    good = p.read_text()
    p.write_text("1 2 3\n")                                  # too few fields
    r = run(["--print-states", str(f)])
    assert r.returncode == 1 and "malformed" in r.stderr and "ehot62_state4.txt" in r.stderr
    p.write_text("0" + good[good.index(" "):])               # padded index 0: a halo site
    r = run(["--print-states", str(f)])
    assert r.returncode == 1 and "outside its block" in r.stderr
    p.write_text(good)
    (d / "ehot62_state9.txt").unlink()
    r = run(["--print-states", str(f)])
    assert r.returncode == 1 and "missing" in r.stderr
    r = run(["--print-states", str(tmp_path / "nope.txt")])
    assert r.returncode == 1
