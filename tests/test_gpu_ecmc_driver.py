"""End-to-end run of the gauge_ecmc_cb driver on the GPU, replayed on the CPU.

The driver keeps the reference ECMC driver's input file, stdout lines, output
files, chain-state files and resume rule (src/bin/gauge_ecmc_ncb.cpp:20-300).
The host mt19937_64 streams drive the hot start (one generator per rank of the
reference's n_core^4 Cartesian grid, seeded seed + rank) and the random shifts
(four uniform_int draws in [0, L_core - 1] each); the chains draw from Philox
counters keyed by splitmix64(seed).  The replay below rebuilds the whole run
with the oracle from those streams -- hot start, shifts, translates, one ECMC
sample per shift on every block, plaquettes, event counts, flows -- and checks
the printed shifts, the plaquette, event-count and topology files, the ILDG
configuration and the per-rank chain-state files; then it runs the driver
again, which resumes from the ILDG file, the chain states and the saved
generator state, and checks the appended outputs the same way.
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
DRIVER = os.path.join(ROOT, "ecmc-qncb_amd", "gauge_ecmc_cb")

# fixed budgets with no common multiple (a tie between two events is decided
# by rounding, which FP64 on two machines need not share)
P = dict(L_core=4, n_core=2, seed=4242, N_therm=1, N_shift=3, beta=6.0, ts=12.3, trs=5.1, trR=3.7,
         N_shift_plaquette=1, N_shift_topo=2, N_steps_gf=3, N_rk_steps=2)
M64 = 2 ** 64 - 1


def write_input(path, run_dir):
    text = f"""L_core = {P['L_core']}
n_core_dims = {P['n_core']}
cold_start = false
seed = {P['seed']}
N_shift_therm = {P['N_therm']}
N_shift = {P['N_shift']}
beta = {P['beta']}
param_theta_sample = {P['ts']}
param_theta_refresh_site = {P['trs']}
param_theta_refresh_R = {P['trR']}
poisson = false
epsilon_set = 0.15
N_shift_plaquette = {P['N_shift_plaquette']}
topo = true
N_shift_topo = {P['N_shift_topo']}
N_steps_gf = {P['N_steps_gf']}
N_rk_steps = {P['N_rk_steps']}
run_name = ec
run_dir = {run_dir}
save_each_shifts = 1000
"""
    with open(path, "w") as f:
        f.write(text)


def chain_key(seed):
    z = (seed + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def rank_block(r, n):
    """block coordinates (x, y, z, t) of Cartesian rank r, row-major with t fastest"""
    return [(r // n ** (3 - d)) % n for d in range(4)]


def hot_start(mt):
    Lc, n = P["L_core"], P["n_core"]
    L = Lc * n
    U = np.zeros((L ** 4, 4, 3, 3), dtype=np.complex128)
    for r in range(n ** 4):
        gen = mt if r == 0 else qo.Mt64(P["seed"] + r)
        b = rank_block(r, n)
        for t in range(Lc):
            for z in range(Lc):
                for y in range(Lc):
                    for x in range(Lc):
                        X, Y, Z, T = b[0] * Lc + x, b[1] * Lc + y, b[2] * Lc + z, b[3] * Lc + t
                        site = ((T * L + Z) * L + Y) * L + X
                        for mu in range(4):
                            U[site, mu] = gen.random_su3()
    return U


def reset(ch, revs):
    for c in ch:
        c.events = c.lifts = 0
        if revs:
            c.revs = 0


def replay(o, mt, ch, params):
    Lc, n = P["L_core"], len(ch)
    shifts, plaq, tqe, nbev = [], [], [], []
    for phase, count in (("therm", P["N_therm"]), ("sample", P["N_shift"])):
        for i in range(count):
            s = [mt.uniform_int(0, Lc - 1) for _ in range(4)]
            shifts.append(s)
            o.translate(s)
            assert o.ecmc_philox(params, ch, chain_key(P["seed"])) == n
            if phase == "therm":
                reset(ch, revs=False)
                continue
            if i % P["N_shift_plaquette"] == 0:
                plaq.append(o.plaquette())
                g_l, g_e, g_r = (sum(getattr(c, f) for c in ch) for f in ("lifts", "events", "revs"))
                nbev.append((g_e // n, g_l // n, g_r // n, P["N_shift_plaquette"] * P["ts"] * n / g_l))
                reset(ch, revs=True)
            if i % P["N_shift_topo"] == 0:
                tqe.extend(o.topo_flowed(0.02, P["N_steps_gf"], P["N_rk_steps"]).tolist())
    return shifts, plaq, tqe, nbev


def run_driver(inp, log):
    assert os.path.exists(DRIVER), "build the driver first (make -C ecmc-qncb_amd)"
    r = subprocess.run([DRIVER, inp], capture_output=True, text=True, timeout=120)
    with open(log, "w") as f:
        f.write(r.stdout + r.stderr)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def read_state(path):
    lines = open(path).read().split("\n")
    h = lines[0].split()
    st = dict(idx=int(h[0]), mu=int(h[1]), eps=int(h[2]), par_refresh_site=float(h[3]),
              par_refresh_R=float(h[4]), th_refresh_site=float(h[5]), th_refresh_R=float(h[6]),
              set_counter=int(h[7]), events=int(h[8]), lifts=int(h[9]), revs=int(h[10]), initialized=int(h[11]))
    st["R"] = np.array([float(v) for ln in lines[1:4] for v in ln.split()])
    st["rng_n"], st["set_gen"] = (int(v) for v in lines[4].split())
    return st


def check_states(state_dir, ch):
    Lc, n = P["L_core"], P["n_core"]
    L, Le = Lc * n, Lc + 2
    by_block = {}
    for c in ch:
        x = [(c.site // L ** d) % L for d in range(4)]
        by_block[tuple(v // Lc for v in x)] = (c, x)
    assert len(by_block) == len(ch)
    for r in range(n ** 4):
        st = read_state(os.path.join(state_dir, f"ec_state{r}.txt"))
        c, x = by_block[tuple(rank_block(r, n))]
        loc = [v % Lc + 1 for v in x]
        assert st["idx"] == ((loc[3] * Le + loc[2]) * Le + loc[1]) * Le + loc[0]
        assert (st["mu"], st["eps"], st["initialized"]) == (c.mu, c.eps, c.initialized)
        assert (st["set_counter"], st["set_gen"], st["rng_n"]) == (c.set_counter, c.set_gen, c.rng_n)
        assert (st["events"], st["lifts"], st["revs"]) == (c.events, c.lifts, c.revs)
        assert np.abs(st["R"] - np.array(c.R[:])).max() < 1e-9
        for f in ("par_refresh_site", "par_refresh_R", "th_refresh_site", "th_refresh_R"):
            assert abs(st[f] - getattr(c, f)) < 1e-8


def glue_state_files(state_dir):
    """rewrite every state file as the reference writes it: line 1 with
    set_counter and events glued (src/io/io.cpp:172-175) and no fifth line,
    so the chains' Philox streams restart from position 0"""
    for name in os.listdir(state_dir):
        path = os.path.join(state_dir, name)
        lines = open(path).read().split("\n")
        h = lines[0].split()
        h[7:9] = [h[7] + h[8]]
        with open(path, "w") as f:
            f.write(" ".join(h) + "\n" + "\n".join(lines[1:4]) + "\n")


def test_ecmc_driver_run_and_resume_match_replay(tmp_path):
    run_dir = str(tmp_path / "runs")
    inp = str(tmp_path / "input.txt")
    write_input(inp, run_dir)
    base = os.path.join(run_dir, "ec", "ec")
    dims = (P["L_core"] * P["n_core"],) * 4

    mt = qo.Mt64(P["seed"])
    o = qo.Lattice(dims)
    o.set_links(hot_start(mt))
    params = qo.ecmc_params(beta=P["beta"], theta_sample=P["ts"], theta_refresh_site=P["trs"],
                            theta_refresh_R=P["trR"], block=(P["L_core"],) * 4)
    ch = o.ecmc_chains(params)
    assert len(ch) == P["n_core"] ** 4

    got_plaq, got_tqe, got_nb = [], [], []
    state_dir = os.path.join(run_dir, "ec", "ec_state")
    # run 1 resumes from the saved run; run 2 resumes from state files in the
    # reference's own layout (set_counter and events glued, no Philox line)
    for run in range(3):
        if run == 2:
            glue_state_files(state_dir)
            for c in ch:
                c.set_counter = c.events = c.rng_n = c.set_gen = 0
        out = run_driver(inp, str(tmp_path / f"run{run}.log"))
        if run >= 1:
            # the resumed driver replaces its hot start by the checkpoint and its
            # generator by the saved one, which is where the replay's stream stands
            assert "Successfully loaded ECMC chain state" in out
            o.set_links(saved)
        shifts, plaq, tqe, nbev = replay(o, mt, ch, params)
        assert [[int(v) for v in m] for m in re.findall(r"Shift \((\d+), (\d+), (\d+), (\d+)\)", out)] == shifts
        got_plaq += plaq
        got_tqe += tqe
        got_nb += nbev
        file_plaq = np.loadtxt(base + "_plaquette.txt")
        assert file_plaq.shape == (len(got_plaq),)
        assert np.abs(file_plaq - np.array(got_plaq)).max() < 2e-10
        file_tqe = np.loadtxt(base + "_topo.txt")
        want = np.array(got_tqe)
        assert file_tqe.shape == want.shape
        assert np.abs(file_tqe[:, 0] - want[:, 0]).max() < 1e-10
        assert np.abs(file_tqe[:, 1] - want[:, 1]).max() < 1e-9
        assert np.abs(file_tqe[:, 2] - want[:, 2]).max() < 1e-9 * max(1.0, np.abs(want[:, 2]).max())
        nb = np.loadtxt(base + "_nbevent.txt", ndmin=2)
        want = np.array(got_nb)
        assert nb.shape == want.shape
        assert np.array_equal(nb[:, :3], want[:, :3])
        assert np.abs(nb[:, 3] / want[:, 3] - 1).max() < 1e-5    # default 6-digit stream output
        check_states(state_dir, ch)
        ctx = qcb.Context(0)
        g = qcb.Lattice(ctx, dims)
        g.ildg_load(base)
        saved = g.get_links()
        assert np.abs(saved - o.links).max() < 1e-9
        g.close()
        ctx.close()
        with open(base + "_params.txt") as f:
            assert f.read().count("Saved final state !") == run + 1
