"""GPU physics checks that pin the device path by what the mathematics fixes:
gauge invariance and covariance, the flow's smoothing and third-order RK3
error, FP32 against FP64 on identical inputs, the heatbath's equilibrium
plaquette against the reference-method chain, and properties at the full
32^4 size.

Reference: gauge transform src/gauge/GaugeField.cpp:100-114; flow
src/flow/gradient_flow.cpp:10-133; observables
src/observables/observables_mpi.cpp:8-178; heatbath
src/heatbath/heatbath_mpi.cpp:11-69.
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as qo
import qcb

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def ctx():
    c = qcb.Context(0)
    yield c
    c.close()


def random_su3(rng, n):
    A = rng.normal(size=(n, 3, 3)) + 1j * rng.normal(size=(n, 3, 3))
    Q, R = np.linalg.qr(A)
    d = np.diagonal(R, axis1=1, axis2=2)
    Q = Q * (d / np.abs(d))[:, None, :]
    det = np.linalg.det(Q)
    return Q / (det ** (1.0 / 3.0))[:, None, None]


def thermal_links(ctx, dims, sweeps=5, seed=21):
    g = qcb.Lattice(ctx, dims)
    g.hot(seed)
    g.hb_sweeps(6.0, 2, sweeps, seed=seed)
    return g


def gauge_rotated(links, dims, G):
    o = qo.Lattice(dims)
    o.set_links(links)
    o.gauge_transform(G)
    return o.links.copy()


def test_gauge_invariance_of_observables(ctx):
    """P, Q and E of the device kernels are invariant under a random gauge
    transform; staples transform as Sigma(x,mu) -> G(x+mu) Sigma G(x)^dag."""
    dims = (8, 4, 6, 8)
    g = thermal_links(ctx, dims)
    rng = np.random.default_rng(5)
    G = random_su3(rng, g.V)
    h = qcb.Lattice(ctx, dims)
    h.set_links(gauge_rotated(g.get_links(), dims, G))
    P0, P1 = g.plaquette(), h.plaquette()
    (Q0, E0), (Q1, E1) = g.clover_qe(), h.clover_qe()
    assert abs(P1 - P0) < 1e-13
    assert abs(E1 - E0) < 1e-12 * E0
    assert abs(Q1 - Q0) < 1e-12 * E0 * g.V
    # covariance of the staples, checked through the plaquette-sum identity
    S0, S1 = g.staples(), h.staples()
    U1 = h.get_links()
    tr0 = np.einsum("smij,smji->sm", g.get_links(), S0)
    tr1 = np.einsum("smij,smji->sm", U1, S1)
    assert np.abs(tr1 - tr0).max() < 1e-12 * np.abs(tr0).max()
    # FP32 kernels on the same pair of fields
    f0, f1 = qcb.Lattice(ctx, dims, qcb.F32), qcb.Lattice(ctx, dims, qcb.F32)
    f0.set_links(g.get_links())
    f1.set_links(U1)
    assert abs(f1.plaquette() - f0.plaquette()) < 2e-6
    assert abs(f1.clover_qe()[1] - f0.clover_qe()[1]) < 1e-5 * E0


def test_flow_gauge_covariance(ctx):
    """Flowing a gauge-rotated field gives the rotated flowed field: P, Q, E
    along the flow agree, and the flowed links are G W G^dag."""
    dims = (4, 6, 4, 8)
    g = thermal_links(ctx, dims, seed=8)
    rng = np.random.default_rng(6)
    G = random_su3(rng, g.V)
    h = qcb.Lattice(ctx, dims)
    h.set_links(gauge_rotated(g.get_links(), dims, G))
    a = g.topo_charge_flowed(0.02, 10, 2)
    b = h.topo_charge_flowed(0.02, 10, 2)
    assert np.abs(a[:, 2] - b[:, 2]).max() < 1e-12 * a[0, 2]
    assert np.abs(a[:, 1] - b[:, 1]).max() < 1e-12 * a[0, 2] * g.V
    g.flow_copy()
    h.flow_copy()
    g.flow_rk3(0.02, 5)
    h.flow_rk3(0.02, 5)
    want = gauge_rotated(g.flow_links(), dims, G)
    assert np.abs(h.flow_links() - want).max() < 1e-12


@pytest.mark.parametrize("prec", [qcb.F64, qcb.F32])
def test_flow_smooths_and_is_third_order(ctx, prec):
    """The clover energy E(t) falls monotonically along the flow; the RK3
    integrator's global error at fixed t is O(eps^3)."""
    dims = (8, 8, 8, 8)
    g = thermal_links(ctx, dims, sweeps=20, seed=3)
    f = qcb.Lattice(ctx, dims, prec)
    f.set_links(g.get_links())
    tqe = f.topo_charge_flowed(0.02, 30, 1)
    assert np.all(np.diff(tqe[:, 2]) < 0)
    assert np.allclose(tqe[:, 0], 0.02 * np.arange(30), atol=1e-13)
    if prec == qcb.F32:
        return
    T = 0.4
    finals = {}
    for n in (5, 10, 20, 80):
        h = qcb.Lattice(ctx, dims)
        h.set_links(g.get_links())
        h.flow_copy()
        h.flow_rk3(T / n, n)
        finals[n] = h.flow_links()
    e = [np.abs(finals[n] - finals[80]).max() for n in (5, 10, 20)]
    r1, r2 = e[0] / e[1], e[1] / e[2]
    assert 6.0 < r1 < 10.5 and 6.0 < r2 < 10.5, (e, r1, r2)


def test_fp32_against_fp64_same_inputs(ctx):
    """The FP32 path against the FP64 path on identical inputs: the plaquette
    of the same links to 1e-6, and a flow to t = 1 with |dQ| < 1e-4 and
    |dE|/E < 1e-5.  One heatbath sweep on identical Philox uniforms is held
    to 1e-5 in <P>: a Kennedy-Pendleton acceptance that rounds the other way
    in FP32 replaces one SU(2) draw, which moves <P> by O(1/V)."""
    dims = (16, 16, 16, 16)
    d = qcb.Lattice(ctx, dims)
    d.hot(11)
    d.hb_sweeps(6.0, 2, 40, seed=11)
    s = qcb.Lattice(ctx, dims, qcb.F32)
    s.set_links(d.get_links())
    assert abs(s.plaquette() - d.plaquette()) < 1e-6
    s.hb_sweeps(6.0, 2, 1, seed=11, sweep_idx=40)
    d.hb_sweeps(6.0, 2, 1, seed=11, sweep_idx=40)
    print(f"one sweep on identical uniforms: |dP| = {abs(s.plaquette() - d.plaquette()):.2e}")
    assert abs(s.plaquette() - d.plaquette()) < 1e-5
    s.set_links(d.get_links())
    a = d.topo_charge_flowed(0.02, 51, 1)
    b = s.topo_charge_flowed(0.02, 51, 1)
    assert abs(a[50, 0] - 1.0) < 1e-12 and abs(b[50, 0] - 1.0) < 1e-12
    assert np.abs(a[:, 1] - b[:, 1]).max() < 1e-4
    assert (np.abs(a[:, 2] - b[:, 2]) / a[:, 2]).max() < 1e-5


@pytest.mark.parametrize("dims", [(8, 8, 8, 8), (64, 4, 4, 4), (12, 6, 6, 8), (24, 4, 4, 6), (6, 4, 4, 4),
                                  (14, 4, 4, 6)])
def test_fp32_flow_links_against_fp64(ctx, dims):
    """The FP32 flow stage takes the lower staples of the planes (mu, 0),
    mu != 0, from the block's shared memory when a block holds whole x rows:
    32 sites when Lx/2 divides 32 (8^4 and 64x4^3), 30 sites with two idle
    lanes for Lx/2 = 6 and 3 (Lx = 12, 6); with Lx/2 = 12 or 7 (Lx = 24, 14)
    its 32 sites are partial rows and every lower staple comes from the field.
    All paths against the FP64 flow of the same links: every link after 10
    RK3 steps at eps = 0.02 within 1e-5 (float rounding; a wrong staple would
    be off by O(eps) ~ 1e-2)."""
    d = thermal_links(ctx, dims, sweeps=3, seed=5)
    s = qcb.Lattice(ctx, dims, qcb.F32)
    s.set_links(d.get_links())
    for g in (d, s):
        g.flow_copy()
        g.flow_rk3(0.02, 10)
    err = np.abs(d.flow_links() - s.flow_links()).max()
    assert err < 1e-5, (dims, err)
    d.close()
    s.close()


@pytest.mark.parametrize("dims", [(8, 8, 8, 8), (16, 8, 8, 8), (64, 4, 4, 4), (12, 6, 6, 8), (24, 4, 4, 6)])
def test_fp32_clover_against_fp64(ctx, dims):
    """The FP32 clover forms leaves 2 and 3 of the planes (0, nu) as the
    x-neighbour's leaves 1 and 4 conjugated by U_0(x - e0) when a wavefront's
    32 sites are whole x rows (Lx/2 divides 32: Lx = 8, 16, 64), and all four
    leaves from the field otherwise (Lx = 12, 24).  Both against the FP64
    clover of the same links, on the gauge field and on a flowed field:
    |dQ| < 2e-6 sqrt(V) and |dE|/E < 1e-6 (float rounding; a wrong leaf
    would be off at O(1))."""
    d = thermal_links(ctx, dims, sweeps=3, seed=9)
    s = qcb.Lattice(ctx, dims, qcb.F32)
    s.set_links(d.get_links())
    V = float(np.prod(dims))
    for step in range(2):
        if step:
            for g in (d, s):
                g.flow_copy()
                g.flow_rk3(0.02, 5)
        q64, e64 = d.flow_qe() if step else d.clover_qe()
        q32, e32 = s.flow_qe() if step else s.clover_qe()
        assert abs(q32 - q64) < 2e-6 * np.sqrt(V), (dims, step, q32, q64)
        assert abs(e32 - e64) / e64 < 1e-6, (dims, step, e32, e64)
    d.close()
    s.close()


def _binned(x, nbin=20):
    x = np.asarray(x)
    bins = x[: len(x) // nbin * nbin].reshape(nbin, -1).mean(axis=1)
    return bins.mean(), bins.std(ddof=1) / np.sqrt(nbin)


@pytest.mark.parametrize("prec", [qcb.F64, qcb.F32])
def test_equilibrium_plaquette_matches_reference_chain(ctx, prec):
    """A free-running device chain (checkerboard order, Philox) and the
    reference-method chain (frozen links, random shifts, lexicographic order,
    mt19937_64; tests/golden/hb_plaquette_8x4.json) sample the same
    equilibrium: their mean plaquettes agree within 3 combined standard errors."""
    with open(os.path.join(HERE, "golden", "hb_plaquette_8x4.json")) as fh:
        gold = json.load(fh)
    dims = tuple(gold["lattice"])
    g = qcb.Lattice(ctx, dims, prec)
    g.hot(7)
    g.hb_sweeps(gold["beta"], gold["n_hits"], 200, seed=2024)
    p = []
    for s in range(4000):
        g.hb_sweeps(gold["beta"], gold["n_hits"], 1, seed=2024, sweep_idx=200 + s)
        if prec == qcb.F32 and s % 4 == 3:
            g.reunitarize()
        p.append(g.plaquette())
    m, e = _binned(p)
    sigma = np.hypot(e, gold["err"])
    print(f"device <P> = {m:.6f} +- {e:.6f}, reference chain {gold['mean']:.6f} +- {gold['err']:.6f}, "
          f"{abs(m - gold['mean']) / sigma:.2f} sigma")
    assert abs(m - gold["mean"]) < 3 * sigma


def test_full_size_32x4_properties(ctx):
    """At the 32^4 bench size: a heatbath chain is reproducible bit for bit,
    stays in SU(3), the FP32 links track the FP64 chain on identical uniforms
    for the first sweep, and the flow of a thermalised field lowers E at every
    step."""
    dims = (32, 32, 32, 32)
    a = qcb.Lattice(ctx, dims, qcb.F32)
    b = qcb.Lattice(ctx, dims, qcb.F32)
    for l in (a, b):
        l.hot(11)
        l.hb_sweeps(6.0, 2, 3, seed=11)
    La, Lb = a.get_links(), b.get_links()
    assert np.array_equal(La, Lb)
    del b
    U = La.reshape(-1, 3, 3)[:: 997]
    assert np.abs(U @ U.conj().transpose(0, 2, 1) - np.eye(3)).max() < 1e-5
    assert np.abs(np.linalg.det(U) - 1).max() < 1e-5
    d = qcb.Lattice(ctx, dims)
    d.hot(11)
    d.hb_sweeps(6.0, 2, 1, seed=11)
    c = qcb.Lattice(ctx, dims, qcb.F32)
    c.hot(11)
    c.hb_sweeps(6.0, 2, 1, seed=11)
    print(f"32^4 one sweep FP32 vs FP64: |dP| = {abs(c.plaquette() - d.plaquette()):.2e}")
    assert abs(c.plaquette() - d.plaquette()) < 1e-5
    del c, d
    # on a thermalised field the clover energy falls at every flow step (on a
    # field only three sweeps from a random start it may first rise: E_clover
    # is not the Wilson action the flow descends)
    a.hb_sweeps(6.0, 2, 30, seed=11, sweep_idx=3)
    tqe = a.topo_charge_flowed(0.02, 6, 1)
    assert np.all(np.diff(tqe[:, 2]) < 0)
    assert len({a.plaquette() for _ in range(3)}) == 1   # fixed-order reduction
