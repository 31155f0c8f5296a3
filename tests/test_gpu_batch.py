"""Batched lattices: every lattice of a batch gives bitwise the results of the
same lattice run alone (blocks never span lattices, per-lattice Philox keys,
fixed-order per-lattice reductions), whatever the launch schedule; graph
replay continues the Philox sweep counter."""
import numpy as np
import pytest

import qcb

pytestmark = pytest.mark.gpu

# the flow stage's blocks: 32 sites in whole rows (Lx = 4, 8), 30 sites in
# whole rows (Lx = 6, 10, 20) and 32 sites in partial rows (Lx = 14), all in
# one launch
DIMS = [(4, 4, 4, 4), (6, 6, 6, 6), (8, 4, 6, 4), (4, 4, 4, 8), (10, 10, 10, 10), (14, 4, 4, 6), (20, 4, 4, 4)]
BETAS = [5.8, 6.0, 6.2, 5.9, 6.1, 6.0, 5.95]
SEEDS = [3, 5, 7, 11, 13, 17, 19]


@pytest.fixture(scope="module")
def ctx():
    c = qcb.Context(0)
    yield c
    c.close()


def alone(ctx, i, prec, n_sweeps, first=0):
    l = qcb.Lattice(ctx, DIMS[i], prec)
    l.hot(SEEDS[i])
    l.hb_sweeps(BETAS[i], 2, n_sweeps, seed=SEEDS[i], sweep_idx=first)
    return l


@pytest.mark.parametrize("prec", [qcb.F64, qcb.F32])
def test_batch_equals_alone(ctx, prec):
    b = qcb.Batch(ctx, DIMS, prec, BETAS, SEEDS)
    b.hot()
    b.hb_sweeps(2, 3)
    P = b.plaquette()
    flows = b.flow(0.02, 4)
    for i in range(len(DIMS)):
        a = alone(ctx, i, prec, 3)
        assert np.array_equal(b.lattice(i).get_links(), a.get_links())
        assert P[i] == a.plaquette()
        t, fin = a.topo_charge_flowed(0.02, 4, 1, final=True)
        assert np.array_equal(flows[i, :4], t)
        assert np.array_equal(flows[i, 4], fin)
        a.close()
    b.close()


def test_persistent_grid_same_chain(ctx):
    b1 = qcb.Batch(ctx, DIMS, qcb.F32, BETAS, SEEDS)
    b2 = qcb.Batch(ctx, DIMS, qcb.F32, BETAS, SEEDS)
    b2.set_schedule(list(range(8)), 64)
    for b in (b1, b2):
        b.hot()
        b.hb_sweeps(2, 2)
    for i in range(len(DIMS)):
        assert np.array_equal(b1.lattice(i).get_links(), b2.lattice(i).get_links())


def test_parity_major_schedule_is_a_valid_sweep(ctx):
    b = qcb.Batch(ctx, DIMS, qcb.F64, BETAS, SEEDS)
    b.set_schedule([0, 2, 4, 6, 1, 3, 5, 7], 0)
    assert b.get_schedule() == ([0, 2, 4, 6, 1, 3, 5, 7], 0)
    b.hot()
    b.hb_sweeps(2, 4)
    for i in range(len(DIMS)):
        U = b.lattice(i).get_links().reshape(-1, 3, 3)
        assert np.abs(U @ U.conj().transpose(0, 2, 1) - np.eye(3)).max() < 1e-12
    assert np.all(b.plaquette() > 0.3)


def test_graph_replay_continues_counter(ctx):
    b = qcb.Batch(ctx, DIMS, qcb.F32, BETAS, SEEDS)
    ref = qcb.Batch(ctx, DIMS, qcb.F32, BETAS, SEEDS)
    for x in (b, ref):
        x.hot()
    b.graph_capture(2, 2, first_sweep=5)
    # one replay at a time, with eager work and synchronisation between
    # replays, and every replay checked against the eager sweeps
    for r in range(3):
        b.graph_launch(1)
        ref.hb_sweeps(2, 2, sweep_idx=5 + 2 * r)
        assert np.array_equal(b.plaquette(), ref.plaquette())
        for i in range(len(DIMS)):
            assert np.array_equal(b.lattice(i).get_links(), ref.lattice(i).get_links())
    b.graph_launch(2)
    ref.hb_sweeps(2, 4, sweep_idx=11)
    for i in range(len(DIMS)):
        assert np.array_equal(b.lattice(i).get_links(), ref.lattice(i).get_links())


def test_autotune_keeps_links(ctx):
    b = qcb.Batch(ctx, DIMS, qcb.F32, BETAS, SEEDS)
    b.hot()
    before = [b.lattice(i).get_links() for i in range(len(DIMS))]
    rep = b.autotune(2, 1)
    assert "ms_per_sweep" in rep
    order, grid = b.get_schedule()
    assert sorted(order) == list(range(8)) and grid % 8 == 0
    for i in range(len(DIMS)):
        assert np.array_equal(b.lattice(i).get_links(), before[i])


@pytest.mark.parametrize("prec", [qcb.F64, qcb.F32])
def test_batch_flow_steps_equal_alone(ctx, prec):
    """qcb_batch_flow_steps leaves every lattice's links bitwise as
    qcb_flow_steps on that lattice alone, and its flow state measurable."""
    b = qcb.Batch(ctx, DIMS, prec, BETAS, SEEDS)
    b.hot()
    b.hb_sweeps(2, 2)
    start = [b.lattice(i).get_links() for i in range(len(DIMS))]
    b.flow_steps(0.02, 3)
    for i, d in enumerate(DIMS):
        g = qcb.Lattice(ctx, d, prec)
        g.set_links(start[i])
        g.flow_steps(0.02, 3)
        assert np.array_equal(b.lattice(i).get_links(), g.get_links())
        assert b.lattice(i).flow_qe() == g.flow_qe()
        g.close()
    b.close()


@pytest.mark.parametrize("prec", [qcb.F64, qcb.F32])
def test_flow_graph_replays_equal_alone(ctx, prec):
    """The batched flow is captured at its first call and replayed on later
    calls with the same eps and step count: every replay, with heatbath sweeps
    and synchronisation between, gives bitwise the single-lattice flow of the
    links it starts from; a new eps or step count captures anew."""
    b = qcb.Batch(ctx, DIMS, prec, BETAS, SEEDS)
    b.hot()
    for r, (eps, n) in enumerate([(0.02, 3), (0.02, 3), (0.02, 3), (0.01, 2), (0.02, 3)]):
        b.hb_sweeps(2, 1, sweep_idx=r)
        start = [b.lattice(i).get_links() for i in range(len(DIMS))]
        flows = b.flow(eps, n)
        for i, d in enumerate(DIMS):
            g = qcb.Lattice(ctx, d, prec)
            g.set_links(start[i])
            t, fin = g.topo_charge_flowed(eps, n, 1, final=True)
            assert np.array_equal(flows[i, :n], t), (r, i)
            assert np.array_equal(flows[i, n], fin), (r, i)
            g.close()
    for r in range(3):
        b.hb_sweeps(2, 1, sweep_idx=10 + r)
        start = [b.lattice(i).get_links() for i in range(len(DIMS))]
        b.flow_steps(0.02, 2)
        for i, d in enumerate(DIMS):
            g = qcb.Lattice(ctx, d, prec)
            g.set_links(start[i])
            g.flow_steps(0.02, 2)
            assert np.array_equal(b.lattice(i).get_links(), g.get_links()), (r, i)
            assert b.lattice(i).flow_qe() == g.flow_qe()
            g.close()
    b.close()


def _config4():
    """64 lattices, 16 each of L = 8, 10, 12, 16 (L^4, periodic), beta ~ U[5.8, 6.2]
    from seed 3, one Philox key per lattice."""
    dims = [(L,) * 4 for L in (8, 10, 12, 16) for _ in range(16)]
    betas = np.random.default_rng(3).uniform(5.8, 6.2, size=64)
    seeds = [3000 + i for i in range(64)]
    return dims, betas, seeds


def test_config4_workload_equals_alone(ctx):
    """The 64-lattice workload as a user runs it: autotuned schedule, 100
    sweeps replayed from a hipGraph of 10, then a 50-step flow measured at
    every step.  Two lattices of each size, run alone with the same schedule,
    give bitwise the same links and (t, Q, E) series; with the default order
    the batch also equals the single-lattice entry points."""
    dims, betas, seeds = _config4()
    b = qcb.Batch(ctx, dims, qcb.F32, betas, seeds)
    b.hot()
    report = b.autotune(2, 1)
    order, grid = b.get_schedule()
    n_streams, tiles = b.get_layout()
    assert "streams=4" in report and "tiles=2x2x2" in report
    b.graph_capture(2, 10, first_sweep=0)
    b.graph_launch(10)
    flows = b.flow(0.02, 50)
    assert flows.shape == (64, 51, 3)
    for i in (0, 5, 16, 27, 32, 40, 48, 63):
        a = qcb.Batch(ctx, [dims[i]], qcb.F32, [betas[i]], [seeds[i]])
        a.set_schedule(order, 0)
        a.set_layout(1, tiles)
        a.hot()
        a.hb_sweeps(2, 100)
        assert np.array_equal(b.lattice(i).get_links(), a.lattice(0).get_links()), i
        fa = a.flow(0.02, 50)
        assert np.array_equal(flows[i], fa[0]), i
        a.close()
    b.close()
    # default (mu-major) order: batch == qcb_hb_sweeps + qcb_topo_charge_flowed on one lattice
    b = qcb.Batch(ctx, dims, qcb.F32, betas, seeds)
    b.hot()
    b.hb_sweeps(2, 100)
    flows = b.flow(0.02, 50)
    for i in (3, 21, 37, 52):
        g = qcb.Lattice(ctx, dims[i], qcb.F32)
        g.hot(seeds[i])
        g.hb_sweeps(betas[i], 2, 100, seed=seeds[i], sweep_idx=0)
        assert np.array_equal(b.lattice(i).get_links(), g.get_links()), i
        t, fin = g.topo_charge_flowed(0.02, 50, 1, final=True)
        assert np.array_equal(flows[i, :50], t) and np.array_equal(flows[i, 50], fin), i
        g.close()
    b.close()


@pytest.mark.parametrize("prec", [qcb.F32, qcb.F64])
def test_stream_groups_same_chain(ctx, prec):
    """Lattices split over 1, 2 or 4 streams, swept directly or from a graph
    replay, follow bitwise the same chain; tile caps change no link."""
    dims, betas, seeds = _config4()
    sel = list(range(0, 64, 5))
    dims, betas, seeds = [dims[i] for i in sel], [betas[i] for i in sel], [seeds[i] for i in sel]
    ref = qcb.Batch(ctx, dims, prec, betas, seeds)
    ref.hot()
    ref.hb_sweeps(2, 4)
    want = [ref.lattice(i).get_links() for i in range(len(dims))]
    ref.close()
    for n, tiles in ((2, (4, 2, 2)), (4, (2, 2, 2)), (4, (8, 4, 4))):
        b = qcb.Batch(ctx, dims, prec, betas, seeds)
        b.set_layout(n, tiles)
        assert b.get_layout() == (n, tiles)
        b.hot()
        b.hb_sweeps(2, 2)
        b.graph_capture(2, 1, first_sweep=2)
        b.graph_launch(1)
        b.graph_launch(1)
        for i in range(len(dims)):
            assert np.array_equal(b.lattice(i).get_links(), want[i]), (n, tiles, i)
        b.close()
