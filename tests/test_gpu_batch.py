"""Batched lattices: every lattice of a batch gives bitwise the results of the
same lattice run alone (blocks never span lattices, per-lattice Philox keys,
fixed-order per-lattice reductions), whatever the launch schedule; graph
replay continues the Philox sweep counter."""
import numpy as np
import pytest

import qcb

pytestmark = pytest.mark.gpu

DIMS = [(4, 4, 4, 4), (6, 6, 6, 6), (8, 4, 6, 4), (4, 4, 4, 8), (10, 10, 10, 10)]
BETAS = [5.8, 6.0, 6.2, 5.9, 6.1]
SEEDS = [3, 5, 7, 11, 13]


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
