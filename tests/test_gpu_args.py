"""Argument and state checks of the C-ABI on a device: every rejected call
returns its negative code with a message in qcb_last_error, touches no
device memory, and leaves the handle usable.  The reference validates none
of these (its drivers trust their input file, src/io/io.cpp:298-349); the
ABI's rules are its own (include/qcb.h)."""
import numpy as np
import pytest

import oracle_lib as qo
import qcb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = qcb.Context(0)
    yield c
    c.close()


def code_of(fn):
    with pytest.raises(qcb.QcbError) as e:
        fn()
    return e.value.code, str(e.value)


@pytest.mark.parametrize("dims", [(4, 4, 4, 5), (2, 4, 4, 4), (4, 0, 4, 4), (4, 4, -4, 4), (3, 3, 3, 3)])
def test_lattice_extents_rejected(ctx, dims):
    code, msg = code_of(lambda: qcb.Lattice(ctx, dims))
    assert code == qcb.EINVAL and "even and >= 4" in msg


def test_lattice_size_and_precision_rejected(ctx):
    # 4 V links must index in 31 bits: 182^4 * 4 > 2^31 (rejected before any allocation)
    code, msg = code_of(lambda: qcb.Lattice(ctx, (182, 182, 182, 182)))
    assert code == qcb.EINVAL and "too large" in msg
    code, msg = code_of(lambda: qcb.Lattice(ctx, (4, 4, 4, 4), prec=7))
    assert code == qcb.EINVAL and "prec" in msg


def test_rejected_calls_leave_the_lattice_usable(ctx):
    dims = (4, 4, 6, 4)
    o = qo.Lattice(dims)
    o.hot_philox(3)
    for s in range(2):                       # leave the random start first
        o.sweep_philox(6.0, 2, 5, 100 + s)
    g = qcb.Lattice(ctx, dims)
    g.set_links(o.links)
    for bad in (lambda: g.hb_sweeps(0.0, 2, 1, seed=1), lambda: g.hb_sweeps(-6.0, 2, 1, seed=1),
                lambda: g.hb_sweeps(6.0, 0, 1, seed=1), lambda: g.hb_sweeps(6.0, 4096, 1, seed=1),
                lambda: g.hb_sweeps(6.0, 2, -1, seed=1), lambda: g.hb_sweeps(float("nan"), 2, 1, seed=1)):
        code, msg = code_of(bad)
        assert code == qcb.EINVAL and "beta > 0" in msg
    for K in (0, 65):
        u = np.full(8 * 2 * 3 * (4 * K + 2) * (g.V // 2), 0.5)
        code, msg = code_of(lambda: g.hb_sweep_injected(6.0, 2, u, K))
        assert code == qcb.EINVAL and "K <= 64" in msg
    # the flow's state rules: measuring or stepping needs a flow copy first
    code, msg = code_of(lambda: g.flow_rk3(0.02, 1))
    assert code == qcb.ESTATE and "qcb_flow_copy" in msg
    code, msg = code_of(lambda: g.flow_qe())
    assert code == qcb.ESTATE
    code, msg = code_of(lambda: g.flow_links())
    assert code == qcb.ESTATE
    # tilings that do not divide the lattice
    code, msg = code_of(lambda: g.set_tiling(4, 4, 4))
    assert code == qcb.EINVAL and "does not divide" in msg
    # nothing above changed the links, and the handle still computes
    assert np.array_equal(g.get_links(), o.links)
    assert abs(g.plaquette() - o.plaquette()) < 1e-12
    g.hb_sweeps(6.0, 2, 1, seed=77)
    o.sweep_philox(6.0, 2, 77, 0)
    assert np.abs(g.get_links() - o.links).max() < 1e-12 * np.abs(o.links).max()
    g.close()


def test_translate_reduces_shifts_modulo_the_extent(ctx):
    dims = (4, 6, 4, 8)
    o = qo.Lattice(dims)
    o.hot_philox(9)
    g = qcb.Lattice(ctx, dims)
    g.set_links(o.links)
    g.translate((4, -6, 8, 16))           # whole turns: no change
    assert np.array_equal(g.get_links(), o.links)
    g.translate((-1, 7, -3, 9))           # = (3, 1, 1, 1)
    o.translate((3, 1, 1, 1))
    assert np.array_equal(g.get_links(), o.links)
    g.close()


def test_context_destroy_with_live_handles_is_refused(ctx):
    c = qcb.Context(0)
    g = qcb.Lattice(c, (4, 4, 4, 4))
    rc = qcb.lib().qcb_ctx_destroy(c.h)
    assert rc == qcb.ESTATE
    assert b"still alive" in qcb.lib().qcb_last_error(c.h)
    assert abs(g.plaquette() - 1.0) < 1e-15   # the context and its lattice still work
    c.close()                             # closes the lattice first, then the context
    assert not c.h



def test_batch_and_service_arguments_rejected(ctx):
    code, msg = code_of(lambda: qcb.Batch(ctx, [(4, 4, 4, 4), (6, 6, 6, 6)], qcb.F32, [6.0, 0.0], [1, 2]))
    assert code == qcb.EINVAL and "lattice 1: need beta > 0" in msg
    code, msg = code_of(lambda: qcb.Batch(ctx, [(4, 4, 4, 4), (6, 5, 6, 6)], qcb.F32, [6.0, 6.0], [1, 2]))
    assert code == qcb.EINVAL and "even and >= 4" in msg
    s = qcb.Service(ctx, arena_bytes=64 << 20)
    for dims, beta, t_max in (((12, 12, 12, 24), 6.0, float("nan")), ((12, 12, 12, 24), 0.0, 1.0),
                              ((12, 12, 12, 24), 6.0, -1.0), ((12, 12, 12, 24), 6.0, 1e9),
                              ((12, 12, 13, 24), 6.0, 1.0), ((256, 256, 256, 256), 6.0, 1.0)):
        code, _ = code_of(lambda: s.submit(1, dims, beta, 7, t_max))
        assert code == qcb.EINVAL, (dims, beta, t_max)
    # a request larger than the whole arena is refused as such
    code, msg = code_of(lambda: s.submit(2, (24, 24, 24, 48), 6.0, 7, 1.0))
    assert code == qcb.ENOMEM and "arena" in msg
    s.close()
