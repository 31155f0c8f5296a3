"""End-to-end run of the gauge_hb_cb driver on the GPU, replayed on the CPU.

The driver (reference src/bin/gauge_hb_ncb.cpp:18-238) keeps the reference's
input file, stdout lines, output files and resume rule.  Every random number
it uses comes from the host mt19937_64(seed): four uniform_int draws per
random shift (src/mpi/Shift.cpp:163-186) and one draw per heatbath call as the
Philox key of that call's checkerboard sweeps.  The replay below rebuilds the
whole run with the oracle from that one stream -- shifts, translates,
sweeps, plaquettes, flows -- and checks the printed shifts, the plaquette and
topology files (10 decimals) and the final ILDG configuration; then it runs
the driver again on the same input, which resumes from the ILDG file and the
saved generator state, and checks the appended outputs the same way.  The
replay of the resumed run starts from the checkpoint the driver wrote, so the
comparison stays a few sweeps deep: FP64 last-bit differences between the
device and the oracle grow from sweep to sweep, as for any heatbath chain.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as qo
import qcb

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "ecmc-qncb_amd", "gauge_hb_cb")

P = dict(L_core=4, seed=777, N_therm=1, N_shift=2, beta=6.0, N_sweeps=1, N_hits=2, N_shift_plaquette=1,
         N_shift_topo=1, N_steps_gf=5, N_rk_steps=1)


def write_input(path, run_dir):
    text = f"""L_core = {P['L_core']}
n_core_dims = 1
cold_start = true
seed = {P['seed']}
N_shift_therm = {P['N_therm']}
N_shift = {P['N_shift']}
beta = {P['beta']}
N_sweeps = {P['N_sweeps']}
N_hits = {P['N_hits']}
N_shift_plaquette = {P['N_shift_plaquette']}
topo = true
N_shift_topo = {P['N_shift_topo']}
N_steps_gf = {P['N_steps_gf']}
N_rk_steps = {P['N_rk_steps']}
run_name = e2e
run_dir = {run_dir}
save_each_shifts = 1000
precision = fp64
"""
    with open(path, "w") as f:
        f.write(text)


def replay(lat, mt):
    """One driver invocation on the oracle; returns shifts, plaquettes, tQE rows."""
    L = P["L_core"]
    shifts, plaq, tqe = [], [], []
    for phase, n in (("therm", P["N_therm"]), ("sample", P["N_shift"])):
        for i in range(n):
            s = [mt.uniform_int(0, L - 1) for _ in range(4)]
            shifts.append(s)
            lat.translate(s)
            key = mt.next()
            for k in range(P["N_sweeps"]):
                lat.sweep_philox(P["beta"], P["N_hits"], key, k)
            if phase == "sample":
                plaq.append(lat.plaquette())
                if i % P["N_shift_topo"] == 0:
                    tqe.extend(lat.topo_flowed(0.02, P["N_steps_gf"], P["N_rk_steps"]).tolist())
    return shifts, plaq, tqe


def run_driver(inp, log):
    assert os.path.exists(DRIVER), "build the driver first (make -C ecmc-qncb_amd)"
    r = subprocess.run([DRIVER, inp], capture_output=True, text=True, timeout=120)
    with open(log, "w") as f:
        f.write(r.stdout + r.stderr)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def printed_shifts(out):
    return [[int(v) for v in m] for m in re.findall(r"Shift \((\d+), (\d+), (\d+), (\d+)\)", out)]


def test_driver_run_and_resume_match_replay(tmp_path):
    run_dir = str(tmp_path / "runs")
    inp = str(tmp_path / "input.txt")
    write_input(inp, run_dir)
    base = os.path.join(run_dir, "e2e", "e2e")

    L = P["L_core"]
    o = qo.Lattice((L,) * 4)
    o.cold()
    mt = qo.Mt64(P["seed"])

    got_plaq, got_tqe = [], []
    for run in range(2):   # second invocation resumes from the ILDG file and seed file
        out = run_driver(inp, str(tmp_path / f"run{run}.log"))
        shifts, plaq, tqe = replay(o, mt)
        assert printed_shifts(out) == shifts
        got_plaq += plaq
        got_tqe += tqe
        file_plaq = np.loadtxt(base + "_plaquette.txt")
        assert file_plaq.shape == (len(got_plaq),)
        assert np.abs(file_plaq - np.array(got_plaq)).max() < 2e-10
        file_tqe = np.loadtxt(base + "_topo.txt")
        want = np.array(got_tqe)
        assert file_tqe.shape == want.shape
        assert np.abs(file_tqe[:, 0] - want[:, 0]).max() < 1e-10
        assert np.abs(file_tqe[:, 1] - want[:, 1]).max() < 1e-9
        assert np.abs(file_tqe[:, 2] - want[:, 2]).max() < 1e-9 * max(1.0, np.abs(want[:, 2]).max())
        # the checkpoint is the replayed configuration; the next replay starts from it
        ctx = qcb.Context(0)
        g = qcb.Lattice(ctx, (L,) * 4)
        g.ildg_load(base)
        saved = g.get_links()
        assert np.abs(saved - o.links).max() < 1e-10
        o.set_links(saved)
        g.close()
        ctx.close()
        with open(base + "_params.txt") as f:
            params = f.read()
        assert params.count("Saved final state !") == run + 1
        assert os.path.exists(os.path.join(run_dir, "e2e", "e2e_seed", "e2e_seed0.txt"))


def test_fp32_driver_tracks_fp64(tmp_path):
    """precision = fp32 (two-row links, reunitarised every 4 sweeps by
    default): the driver draws the same shifts and Philox keys from the host
    stream as in FP64, so its heatbath consumes identical uniforms and its
    plaquettes track the FP64 run's (two independent 8^4 chains differ by
    ~1e-3); the FP32 ILDG checkpoint holds SU(3) links."""
    outs = {}
    for prec in ("fp64", "fp32"):
        run_dir = tmp_path / prec
        inp = tmp_path / f"{prec}.txt"
        inp.write_text(f"""L_core = 8
n_core_dims = 1
cold_start = true
seed = 4321
N_shift_therm = 0
N_shift = 6
beta = 6.0
N_sweeps = 1
N_hits = 2
N_shift_plaquette = 1
topo = false
run_name = p
run_dir = {run_dir}
save_each_shifts = 1000
precision = {prec}
""")
        out = run_driver(str(inp), str(tmp_path / f"{prec}.log"))
        shifts = re.findall(r"Shift \((\d+), (\d+), (\d+), (\d+)\)", out)
        plaq = np.loadtxt(os.path.join(run_dir, "p", "p_plaquette.txt"))
        outs[prec] = (shifts, plaq, os.path.join(run_dir, "p", "p"))
    s64, p64, _ = outs["fp64"]
    s32, p32, ildg32 = outs["fp32"]
    assert s32 == s64 and len(s64) == 6
    assert p32.shape == p64.shape == (6,)
    # measured: |dP| ~1e-7 on all six shifts
    assert np.abs(p32 - p64).max() < 1e-5
    assert np.all(p32 > 0.2) and np.all(p32 < 0.8)
    ctx = qcb.Context(0)
    g = qcb.Lattice(ctx, (8, 8, 8, 8))
    g.ildg_load(ildg32)
    U = g.get_links().reshape(-1, 3, 3)
    assert np.abs(U @ U.conj().transpose(0, 2, 1) - np.eye(3)).max() < 1e-5
    assert abs(g.plaquette() - p32[-1]) < 1e-5
    g.close()
    ctx.close()
