"""The plain FP64 configuration at its own size against the FP64 oracle.

16^4 periodic lattice, hot start with Haar SU(3) links from seed 7 (complex
Gaussian 3x3, QR, phase-fixed to det = 1; the device's Philox start and the
oracle's restatement of it), beta = 6.0, N_hits = 2, 200 heatbath sweeps, then
a Wilson flow of 100 RK3 steps at eps = 0.02 measuring the plaquette,
Q_clover(t) and E_clover(t) at every step (src/heatbath/heatbath_mpi.cpp:11-69,
src/flow/gradient_flow.cpp:117-128, src/observables/observables_mpi.cpp:8-178).

Tolerances: 1e-12 relative for links after one sweep from identical links
(injected uniforms, and the Philox counters), for P and E at every flow step,
and |dQ| < 1e-12 E V; the links after all 100 flow steps to 1e-11 (the
difference of FMA-contracted and separate arithmetic is carried through 300
stages).  The 200 sweeps themselves are not compared link by link: last-bit
differences grow ~10^3 per sweep in a Monte Carlo chain, so parity is taken
one sweep deep from the GPU's own thermalised links.
"""
import numpy as np
import pytest

import oracle_lib as qo
import qcb

pytestmark = pytest.mark.gpu

DIMS = (16, 16, 16, 16)
REL = 1e-12


def maxrel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def ctx():
    c = qcb.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def thermalised(ctx):
    g = qcb.Lattice(ctx, DIMS, qcb.F64)
    g.hot(7)
    o = qo.Lattice(DIMS)
    o.hot_philox(7)
    assert maxrel(g.get_links(), o.links) < 1e-13
    g.hb_sweeps(6.0, 2, 200, seed=7, sweep_idx=0)
    X = g.get_links()
    g.close()
    return X


def test_sweep_from_thermalised_links(ctx, thermalised):
    X = thermalised
    g = qcb.Lattice(ctx, DIMS, qcb.F64)
    o = qo.Lattice(DIMS)
    # one sweep on injected uniforms: Vh x 8 half-sweeps x 2 hits x 3 subgroups x (4K+2) slots
    K = 8
    u = np.random.default_rng(22).uniform(size=8 * 2 * 3 * (4 * K + 2) * (g.V // 2))
    g.set_links(X)
    o.set_links(X)
    ex_g = g.hb_sweep_injected(6.0, 2, u, K)
    ex_o = o.sweep_injected(6.0, 2, u, K)
    assert ex_g == ex_o
    assert maxrel(g.get_links(), o.links) < REL
    del u
    # sweep 200 of the Philox chain, as the 201st sweep of the run would draw it
    g.set_links(X)
    o.set_links(X)
    g.hb_sweeps(6.0, 2, 1, seed=7, sweep_idx=200)
    o.sweep_philox(6.0, 2, 7, 200)
    assert maxrel(g.get_links(), o.links) < REL
    assert abs(g.plaquette() - o.plaquette()) < REL * o.plaquette()
    g.close()


def test_flow_measurements_every_step(ctx, thermalised):
    X = thermalised
    eps, n = 0.02, 100
    g = qcb.Lattice(ctx, DIMS, qcb.F64)
    g.set_links(X)
    o = qo.Lattice(DIMS)
    o.set_links(X)
    g.flow_copy()
    V = g.V
    for s in range(n + 1):
        Pg = g.flow_plaquette()
        Qg, Eg = g.flow_qe()
        Po = o.plaquette()
        Qo, Eo = o.clover_qe()
        assert abs(Pg - Po) < REL * Po, (s, Pg, Po)
        assert abs(Eg - Eo) < REL * Eo, (s, Eg, Eo)
        assert abs(Qg - Qo) < REL * Eo * V, (s, Qg, Qo)
        if s < n:
            g.flow_rk3(eps, 1)
            o.flow_step(eps)
    assert maxrel(g.flow_links(), o.links) < 1e-11
    # the topo_charge_flowed schedule on the same links gives the same series
    tqe = g.topo_charge_flowed(eps, 11, 2)
    ref = qo.Lattice(DIMS)
    ref.set_links(X)
    want = ref.topo_flowed(eps, 11, 2)
    assert np.array_equal(tqe[:, 0], want[:, 0])
    assert np.abs(tqe[:, 2] - want[:, 2]).max() < REL * want[:, 2].max()
    assert np.abs(tqe[:, 1] - want[:, 1]).max() < REL * want[:, 2].max() * V
    # the flowed field is smooth: E falls monotonically and P rises
    assert np.all(np.diff(tqe[:, 2]) < 0)
    g.close()
