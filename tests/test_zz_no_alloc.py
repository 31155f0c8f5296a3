"""qcb_topo_charge_flowed_noalloc against qcb_topo_charge_flowed: the entry
point that allocates nothing (its measurement results staged in the library's
static device scratch array) gives bitwise the series of the allocating one,
on two calls in a row, the second with a smaller input; its refusals (no flow
buffers yet, too many measurements) leave nothing half done."""
import numpy as np
import pytest

import qcb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = qcb.Context(0)
    yield c
    c.close()


def _pair(ctx, dims, prec, seed):
    a = qcb.Lattice(ctx, dims, prec)
    a.hot(seed)
    a.hb_sweeps(6.0, 2, 3, seed=seed)
    b = qcb.Lattice(ctx, dims, prec)
    b.set_links(a.get_links())
    return a, b


def test_noalloc_equals_allocating_entry_point(ctx):
    # first call: 8^3 x 16 FP64, 40 measurements; second: 4^4 FP32, 6 measurements
    for dims, prec, n_gf, n_rk in (((8, 8, 8, 16), qcb.F64, 40, 2), ((4, 4, 4, 4), qcb.F32, 6, 3)):
        a, b = _pair(ctx, dims, prec, 17)
        want, want_fin = a.topo_charge_flowed(0.02, n_gf, n_rk, final=True)
        with pytest.raises(qcb.QcbError) as e:        # no flow buffers yet: refuses, allocates nothing
            b.topo_charge_flowed_noalloc(0.02, n_gf, n_rk, final=True)
        assert e.value.code == qcb.ESTATE
        b.flow_copy()                                  # set-up: creates the flow buffers
        got, got_fin = b.topo_charge_flowed_noalloc(0.02, n_gf, n_rk, final=True)
        assert np.array_equal(got, want)
        assert np.array_equal(got_fin, want_fin)
        # without the final measurement, and again (the scratch keeps the last call's contents)
        assert np.array_equal(b.topo_charge_flowed_noalloc(0.02, n_gf, n_rk), want)
        assert np.array_equal(b.get_links(), a.get_links())   # the links are not flowed
        a.close()
        b.close()


def test_noalloc_limits(ctx):
    a, b = _pair(ctx, (4, 4, 4, 4), qcb.F64, 3)
    b.flow_copy()
    with pytest.raises(qcb.QcbError) as e:
        b.topo_charge_flowed_noalloc(0.02, 43690, 1)
    assert e.value.code == qcb.EINVAL
    got = b.topo_charge_flowed_noalloc(0.02, 43689, 0)      # the largest count, no steps
    assert got.shape == (43689, 3) and np.all(got[:, 0] == 0.0)
    assert np.all(got[:, 2] == got[0, 2])
    assert np.array_equal(got[:3], a.topo_charge_flowed(0.02, 3, 0))
