"""The flow's link fields are stored in their own order (FP32 entries in
16-byte pairs) and converted on the way in (links -> flow field) and out
(flow field -> links, and the download of a flow field).  These checks pin
the conversions in FP32, where the two orders differ: copying the links into
the flow field and reading it back gives the links bitwise; flowing in place
(qcb_flow_steps) leaves in the links bitwise the flow field that the same
steps leave behind; and the batched flow of lattices whose site counts are not
multiples of 64 equals each lattice alone."""
import numpy as np
import pytest

import qcb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = qcb.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (6, 6, 6, 6), (8, 6, 10, 4)])
def test_flow_field_round_trip(ctx, dims):
    g = qcb.Lattice(ctx, dims, qcb.F32)
    g.hot(5)
    g.hb_sweeps(6.0, 2, 2, seed=5)
    U = g.get_links()
    g.flow_copy()
    assert np.array_equal(g.flow_links(), U)
    g.flow_rk3(0.02, 3)
    W = g.flow_links()
    assert not np.array_equal(W, U)
    h = qcb.Lattice(ctx, dims, qcb.F32)
    h.set_links(U)
    h.flow_steps(0.02, 3)
    assert np.array_equal(h.get_links(), W)
    assert np.array_equal(h.flow_links(), W)
    g.close()
    h.close()


def test_batch_flow_uneven_sizes(ctx):
    dims = [(6, 6, 6, 6), (4, 6, 8, 6), (10, 4, 4, 6)]
    b = qcb.Batch(ctx, dims, qcb.F32, [6.0, 5.9, 6.1], [1, 2, 3])
    b.hot()
    b.hb_sweeps(2, 2)
    start = [b.lattice(i).get_links() for i in range(len(dims))]
    b.flow_steps(0.02, 2)
    for i, d in enumerate(dims):
        g = qcb.Lattice(ctx, d, qcb.F32)
        g.set_links(start[i])
        g.flow_steps(0.02, 2)
        assert np.array_equal(b.lattice(i).get_links(), g.get_links()), i
        g.close()
    b.close()
