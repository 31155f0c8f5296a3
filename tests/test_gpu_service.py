"""Flow-measurement service: a request's (t, Q, E) series is bitwise the same
alone and in a mix of other requests, and equals the single-lattice path
(hot start, sweeps with reunitarisation every 4, topo_charge_flowed)."""
import numpy as np
import pytest

import qcb

pytestmark = pytest.mark.gpu

N_THERM, EPS = 6, 0.02


@pytest.fixture(scope="module")
def ctx():
    c = qcb.Context(0)
    yield c
    c.close()


def run(svc, reqs):
    """reqs: list of (id, dims, beta, seed, t_max, submit_after_iteration)"""
    pending = sorted(reqs, key=lambda r: r[5])
    out, it = {}, 0
    while pending or svc.status()["resident"] or svc.status()["queued"]:
        while pending and pending[0][5] <= it:
            r = pending.pop(0)
            svc.submit(r[0], r[1], r[2], r[3], r[4])
        svc.step()
        it += 1
        for rid, tqe, times in svc.poll():
            out[rid] = (tqe, times)
        assert it < 10000
    return out


@pytest.mark.parametrize("prec", [qcb.F32, qcb.F64])
def test_alone_equals_mixed(ctx, prec):
    target = (7, (8, 8, 8, 16), 6.0, 99, 0.1, 0)
    alone = run(qcb.Service(ctx, prec, 1 << 30, 2, N_THERM, 4, EPS), [target])[7][0]
    mix = [(1, (12, 12, 12, 24), 5.9, 1, 0.14, 0), (2, (4, 4, 4, 8), 6.1, 2, 0.04, 1),
           (7, (8, 8, 8, 16), 6.0, 99, 0.1, 3), (3, (8, 8, 8, 16), 6.0, 98, 0.2, 3),
           (4, (16, 16, 16, 32), 6.0, 4, 0.06, 5)]
    # a small arena forces queueing and first-fit admission
    svc = qcb.Service(ctx, prec, 300 << 20, 2, N_THERM, 4, EPS)
    got = run(svc, mix)
    assert set(got) == {1, 2, 3, 4, 7}
    assert np.array_equal(got[7][0], alone)
    assert alone.shape == (6, 3)
    for rid, (tqe, (ta, tadm, td)) in got.items():
        assert ta <= tadm <= td


def test_service_equals_single_lattice_path(ctx):
    dims, beta, seed, t_max = (8, 8, 8, 16), 6.0, 5, 0.08
    res = run(qcb.Service(ctx, qcb.F32, 1 << 30, 2, N_THERM, 4, EPS), [(1, dims, beta, seed, t_max, 0)])[1][0]
    lat = qcb.Lattice(ctx, dims, qcb.F32)
    lat.hot(seed)
    for s in range(N_THERM):
        lat.hb_sweeps(beta, 2, 1, seed=seed, sweep_idx=s)
        if (s + 1) % 4 == 0:
            lat.reunitarize()
    t, fin = lat.topo_charge_flowed(EPS, 4, 1, final=True)
    assert np.array_equal(res[:4], t)
    assert np.array_equal(res[4], fin)


def test_config5_sizes_alone_equals_mixed(ctx):
    """Requests of the serving workload's sizes, L^3 x 2L with L in {12, 16, 24},
    flowed to t_max in [0.5, 2.0]: the 24^3 x 48 request's series is bitwise
    the same alone and in a mix that arrives while it is resident."""
    target = (9, (24, 24, 24, 48), 6.0, 123, 0.5, 0)
    alone = run(qcb.Service(ctx, qcb.F32, 2 << 30, 2, N_THERM, 4, EPS), [target])[9][0]
    assert alone.shape == (26, 3)
    assert np.allclose(alone[:, 0], EPS * np.arange(26))
    mix = [(1, (12, 12, 12, 24), 5.9, 1, 2.0, 0), (2, (16, 16, 16, 32), 6.1, 2, 1.0, 2),
           (9, (24, 24, 24, 48), 6.0, 123, 0.5, 3), (3, (12, 12, 12, 24), 6.0, 3, 0.5, 4),
           (4, (24, 24, 24, 48), 5.9, 4, 0.8, 5), (5, (16, 16, 16, 32), 6.0, 5, 1.5, 30)]
    # two 24^3 x 48 requests (~490 MB each in FP32) cannot both fit: the second queues
    got = run(qcb.Service(ctx, qcb.F32, 900 << 20, 2, N_THERM, 4, EPS), mix)
    assert set(got) == {1, 2, 3, 4, 5, 9}
    assert np.array_equal(got[9][0], alone)
    assert got[1][0].shape == (101, 3)
    assert got[4][1][1] >= got[9][1][2]          # admitted once request 9's slice was freed


def test_edge_schedules_and_many_requests(ctx):
    """No thermalisation and t_max = 0: a request's first iteration both hot-starts
    and prepares the flow, and a zero-length flow is one measurement.  Many
    small requests at once grow the descriptor sets past their first size."""
    svc = qcb.Service(ctx, qcb.F64, 1 << 30, 2, 0, 0, EPS)
    reqs = [(i, (4, 4, 4, 8), 6.0, 50 + i, [0.0, 0.04, 0.1][i % 3], i % 5) for i in range(90)]
    got = run(svc, reqs)
    assert set(got) == set(range(90))
    for i in (0, 1, 2, 47):
        lat = qcb.Lattice(ctx, (4, 4, 4, 8), qcb.F64)
        lat.hot(50 + i)
        n = int(round(reqs[i][4] / EPS))
        t, fin = lat.topo_charge_flowed(EPS, n, 1, final=True)
        assert np.array_equal(got[i][0][:n], t) and np.array_equal(got[i][0][n], fin), i
        lat.close()
