"""Kennedy-Pendleton attempt budget (Philox heatbath).

The reference draws until acceptance (src/heatbath/heatbath.cpp:14-26); the
device caps each draw at qcb_ctx_set_kp_budget attempts (default 1024), skips
the subgroup update of a draw that runs out, and the sweep call reports it as
QCB_EBUDGET.  The condition is forced from the host so every kernel exits
normally: at beta = 1e-12 no attempt can accept, because the largest uniform
is 1 - 2^-33 (so -ln r1 >= 1.16e-10) while alpha = 2 beta k / 3 < 1e-11 for
any staple (k <= 12), which makes lambda^2 >= 8 > 1 - delta^2.  With a cap of
one attempt every draw is exhausted and the links must not move.
"""
import numpy as np
import pytest

import qcb

pytestmark = pytest.mark.gpu

TINY_BETA = 1e-12


@pytest.fixture()
def ctx():
    c = qcb.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("prec", [qcb.F64, qcb.F32])
def test_single_lattice_reports_exhausted_draws(ctx, prec):
    dims = (4, 4, 4, 6)
    g = qcb.Lattice(ctx, dims, prec)
    g.hot(5)
    before = g.get_links()
    ctx.set_kp_budget(1)
    with pytest.raises(qcb.QcbError) as e:
        g.hb_sweeps(TINY_BETA, 2, 1, seed=9)
    assert e.value.code == qcb.EBUDGET
    n_draws = 4 * g.V * 2 * 3                       # links x hits x subgroups, all exhausted
    assert f"{n_draws} Kennedy-Pendleton draws" in str(e.value)
    after = g.get_links()
    if prec == qcb.F64:
        assert np.array_equal(after, before)
    else:   # FP32 stores re-orthonormalise rows 0 and 1: unchanged up to float rounding
        assert np.abs(after - before).max() < 3e-7
    # the default budget at beta = 6 runs clean, and the counter was reset
    ctx.set_kp_budget(1024)
    g.hb_sweeps(6.0, 2, 3, seed=9)
    with pytest.raises(qcb.QcbError) as e:
        ctx.set_kp_budget(0)
    assert e.value.code == qcb.EINVAL


def test_batch_names_the_first_exhausted_lattice(ctx):
    """Each lattice of a batch counts its own exhausted draws; at beta = 1e-12
    no attempt accepts, so lattices 1 and 2 exhaust every draw whatever the
    cap, while lattice 0 (beta = 6) updates normally."""
    dims = [(4, 4, 4, 4), (6, 4, 4, 4), (4, 6, 4, 4)]
    b = qcb.Batch(ctx, dims, qcb.F32, [6.0, TINY_BETA, TINY_BETA], [1, 2, 3])
    b.hot()
    before = [b.lattice(i).get_links() for i in range(3)]
    with pytest.raises(qcb.QcbError) as e:
        b.hb_sweeps(2, 1)
    assert e.value.code == qcb.EBUDGET
    n_draws = sum(4 * int(np.prod(d)) for d in dims[1:]) * 2 * 3
    assert f"{n_draws} Kennedy-Pendleton draws" in str(e.value) and "first in lattice 1" in str(e.value)
    assert np.abs(b.lattice(0).get_links() - before[0]).max() > 1e-2
    for i in (1, 2):
        assert np.abs(b.lattice(i).get_links() - before[i]).max() < 3e-7
    # a cap of one attempt, and graph replays, report through the same counters
    ctx.set_kp_budget(1)
    with pytest.raises(qcb.QcbError) as e:
        b.hb_sweeps(2, 1, sweep_idx=1)
    assert "first in lattice 0" in str(e.value)
    b.graph_capture(2, 2, first_sweep=3)
    with pytest.raises(qcb.QcbError) as e:
        b.graph_launch(1)
    assert e.value.code == qcb.EBUDGET
    ctx.set_kp_budget(1024)
    b.close()
    b2 = qcb.Batch(ctx, dims[:2], qcb.F32, [6.0, 5.8], [1, 2])
    b2.hot()
    b2.hb_sweeps(2, 2)                              # clean at beta ~ 6
    b2.graph_capture(2, 2, first_sweep=2)
    b2.graph_launch(2)
    b2.close()


def test_service_flags_the_request(ctx):
    ctx.set_kp_budget(1)
    svc = qcb.Service(ctx, qcb.F32, arena_bytes=64 << 20, n_hits=2, n_therm=2, reunit_every=0)
    svc.submit(7, (4, 4, 4, 8), TINY_BETA, 11, 0.04)
    while svc.step():
        pass
    with pytest.raises(qcb.QcbError) as e:
        svc.poll()
    assert e.value.code == qcb.EBUDGET
    rid, tqe, _ = e.value.result
    assert rid == 7 and tqe.shape == (3, 3)
    svc.close()
