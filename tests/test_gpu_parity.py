"""GPU parity: the HIP path through the C-ABI against the FP64 oracle.

Tolerances: FP64 kernels agree with the oracle to 1e-12 relative (the bar
for the deterministic kernels -- staples, plaquette, flow, Q, E -- and for
heatbath sweeps fed injected uniforms); Q is compared in absolute terms scaled by E*V because it is
near zero.  FP32 kernels are held to float-rounding bounds stated per test.
"""
import numpy as np
import pytest

import oracle_lib as qo
import qcb
from linkstore import assert_fp32_stored

pytestmark = pytest.mark.gpu

REL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    c = qcb.Context(0)
    yield c
    c.close()


def hot_pair(ctx, dims, seed, prec=qcb.F64):
    g = qcb.Lattice(ctx, dims, prec)
    g.hot(seed)
    o = qo.Lattice(dims)
    o.hot_philox(seed)
    return g, o


def maxrel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_roundtrip_and_cold(ctx):
    dims = (4, 6, 4, 8)
    g = qcb.Lattice(ctx, dims)
    assert g.plaquette() == 1.0
    assert g.clover_qe() == (0.0, 0.0)
    rng = np.random.default_rng(0)
    U = rng.normal(size=(g.V, 4, 3, 3)) + 1j * rng.normal(size=(g.V, 4, 3, 3))
    g.set_links(U)
    assert np.array_equal(g.get_links(), U)
    f = qcb.Lattice(ctx, dims, qcb.F32)
    f.set_links(U)
    assert_fp32_stored(f.get_links(), U)


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (8, 8, 8, 8), (4, 6, 8, 6)])
def test_hot_start_matches_oracle(ctx, dims):
    g, o = hot_pair(ctx, dims, 7)
    assert maxrel(g.get_links(), o.links) < 1e-13
    U = g.get_links().reshape(-1, 3, 3)
    assert np.abs(U @ U.conj().transpose(0, 2, 1) - np.eye(3)).max() < 1e-13
    assert np.abs(np.linalg.det(U) - 1).max() < 1e-13


# (32, 8, 4, 4): the clover's wavefronts hold two whole x rows and most take
# their y-planes' leaves 2 and 3 from the y-neighbour (8^4: x-planes only)
@pytest.mark.parametrize("dims", [(8, 8, 8, 8), (4, 6, 8, 6), (32, 8, 4, 4)])
def test_plaquette_clover_staples(ctx, dims):
    g, o = hot_pair(ctx, dims, 11)
    # a few oracle sweeps make a non-trivial, partly ordered field
    for s in range(2):
        o.sweep_philox(6.0, 1, 3, s)
    g.set_links(o.links)
    assert abs(g.plaquette() - o.plaquette()) < REL * abs(o.plaquette())
    Qg, Eg = g.clover_qe()
    Qo, Eo = o.clover_qe()
    assert abs(Eg - Eo) < REL * Eo
    assert abs(Qg - Qo) < REL * Eo * g.V
    st = g.staples()
    want = np.stack([[o.staple(s, mu) for mu in range(4)] for s in range(g.V)])
    assert maxrel(st, want) < REL
    # FP32 (two-row links, rows rebuilt inside the staple products): float
    # rounding of six sums of two products
    f = qcb.Lattice(ctx, dims, qcb.F32)
    f.set_links(o.links)
    assert np.abs(f.staples() - want).max() < 2e-5
    assert abs(f.plaquette() - o.plaquette()) < 1e-6
    f.close()


# Lx/2 in {2, 4, 8, 16, 32} (whole x rows in 32-site wavefronts) and
# {3, 5, 6} (whole rows in 30-site wavefronts, whose first rows find their
# y-neighbours past the previous wavefront's idle lanes) with y tiles of 4, 6
# and 8 rows: wavefronts with and without all their y-neighbours in the
# block; Lx = 14 has no whole rows per wavefront and shares nothing
@pytest.mark.parametrize("dims", [(4, 8, 4, 6), (8, 4, 8, 4), (16, 6, 4, 4), (32, 4, 6, 4), (64, 4, 4, 4),
                                  (12, 8, 4, 4), (10, 8, 4, 4), (6, 6, 4, 4), (14, 4, 4, 4)])
def test_clover_shared_leaves_match_oracle(ctx, dims):
    """The clover's shared leaves (x-planes from the x-neighbour, y-planes
    from the y-neighbour) against the oracle's four-leaf clover, FP64 to
    1e-12, on the links and on a flowed field."""
    g, o = hot_pair(ctx, dims, 19)
    for s in range(2):
        o.sweep_philox(6.0, 1, 5, s)
    g.set_links(o.links)
    Qg, Eg = g.clover_qe()
    Qo, Eo = o.clover_qe()
    assert abs(Eg - Eo) < REL * Eo
    assert abs(Qg - Qo) < REL * Eo * g.V
    tg = g.topo_charge_flowed(0.02, 3, 1)
    to = o.topo_flowed(0.02, 3, 1)
    assert np.all(np.abs(tg[:, 2] - to[:, 2]) < REL * to[:, 2])
    assert np.all(np.abs(tg[:, 1] - to[:, 1]) < REL * to[:, 2] * g.V)
    g.close()


def test_abelian_flux_closed_form(ctx):
    from test_oracle_lattice import abelian_field
    dims = (6, 4, 4, 8)
    U, f, gg = abelian_field(dims, (1, -1, 0), (2, -1, -1))
    g = qcb.Lattice(ctx, dims)
    g.set_links(U)
    V = g.V
    P = (np.cos(f).sum() / 3 + np.cos(gg).sum() / 3 + 4) / 6
    assert abs(g.plaquette() - P) < 1e-14
    Q, E = g.clover_qe()
    assert abs(Q + V * np.sum(np.sin(f) * np.sin(gg)) / (4 * np.pi ** 2)) < 1e-12
    assert abs(E - 0.5 * np.sum(np.sin(f) ** 2 + np.sin(gg) ** 2)) < 1e-14


def _su3_algebra(rng, n, scale):
    A = rng.normal(size=(n, 3, 3)) + 1j * rng.normal(size=(n, 3, 3))
    Z = 0.5 * (A - A.conj().transpose(0, 2, 1))
    Z -= np.trace(Z, axis1=1, axis2=2)[:, None, None] / 3 * np.eye(3)
    return scale * Z


def test_probe_exp_proj_reunit(ctx):
    rng = np.random.default_rng(1)
    Z = _su3_algebra(rng, 64, 0.4)
    flat = Z.view(np.float64).reshape(64, 18)
    out = ctx.probe(qcb.F64, qcb.PROBE_EXP, flat, 0.7).view(np.complex128).reshape(64, 3, 3)
    want = np.stack([qo.exp_luscher(z, 0.7) for z in Z])
    assert np.abs(out - want).max() < 1e-14
    out32 = ctx.probe(qcb.F32, qcb.PROBE_EXP, flat, 0.7).view(np.complex128).reshape(64, 3, 3)
    assert np.abs(out32 - want).max() < 2e-6
    A = (rng.normal(size=(64, 3, 3)) + 1j * rng.normal(size=(64, 3, 3)))
    out = ctx.probe(qcb.F64, qcb.PROBE_PROJ, A.view(np.float64).reshape(64, 18)).view(np.complex128).reshape(64, 3, 3)
    assert np.abs(out - np.stack([qo.proj_lie(a) for a in A])).max() < 1e-15
    out = ctx.probe(qcb.F64, qcb.PROBE_REUNIT, A.view(np.float64).reshape(64, 18)).view(np.complex128).reshape(64, 3, 3)
    for a, b in zip(A, out):
        x = qo.m3(a)
        qo.lib().qo_reunitarize_link(qo.dp(x))
        assert np.abs(qo.c3(x) - b).max() < 1e-14


def test_probe_exp_fp32_term_tiers(ctx):
    """The FP32 exponential takes 5, 8 or 12 Taylor terms per wavefront, by
    the largest |X|_F of its lanes (Taylor remainder below 2^-26 for
    |X|_F <= 0.148, 0.556 and ~3).  One wavefront per call, every lane at the
    same norm, so each call runs one tier: |X|_F = 0.05 and 0.14 (5 terms),
    0.3 and 0.55 (8), 1.0 and 2.5 (12).  Each result is unitary to float
    rounding and equals scipy's expm to 2e-6 (5e-6 at |X|_F = 2.5, where
    the coefficient recursion accumulates more rounding)."""
    from scipy.linalg import expm
    rng = np.random.default_rng(5)
    Z = _su3_algebra(rng, 64, 1.0)
    Z /= np.linalg.norm(Z, axis=(1, 2))[:, None, None]
    flat = Z.view(np.float64).reshape(64, 18)
    for norm, tol in ((0.05, 2e-6), (0.14, 2e-6), (0.3, 2e-6), (0.55, 2e-6), (1.0, 2e-6), (2.5, 5e-6)):
        out = ctx.probe(qcb.F32, qcb.PROBE_EXP, flat, norm).view(np.complex128).reshape(64, 3, 3)
        want = np.stack([expm(norm * z) for z in Z])
        assert np.abs(out - want).max() < tol, norm
        uu = np.einsum("nij,nkj->nik", out, out.conj())
        assert np.abs(uu - np.eye(3)).max() < 2e-6, norm


def _kp_inputs(seed, n, K):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.uniform(0.05, 3.0, size=(n, 1)), rng.uniform(size=(n, 4 * K + 2))], 1)


def test_probe_kp_fp64(ctx):
    """Kennedy-Pendleton draws on injected uniforms: accepted attempt index
    exact and x to 1e-14 against the oracle (heatbath.cpp:8-36)."""
    K, n = 6, 512
    inp = _kp_inputs(2, n, K)
    out = ctx.probe(qcb.F64, qcb.PROBE_KP, inp, 6.0, K)
    for row, o in zip(inp, out):
        x = np.zeros(4)
        a = qo.lib().qo_kp_injected(6.0, row[0], qo.dp(np.ascontiguousarray(row[1:])), K, qo.dp(x))
        assert a == int(o[0])
        if a >= 0:
            assert np.abs(o[1:] - x).max() < 1e-14


def test_probe_kp_fp32(ctx):
    """The float Kennedy-Pendleton draw (hardware log/cos/sin/rcp) against
    the FP64 oracle.  Items whose accept/reject decisions are within 1e-5 of
    the boundary delta^2 = 1 - lambda^2 at some attempt up to the accepted one
    are float-sensitive by construction and excluded (at most 1% of them).
    For the rest the accepted attempt is the oracle's and
    |dx0| < 4e-6, |dx_c| < 1e-5 + 4e-7 / r with r = sqrt(1 - a0^2): a0 = 1 -
    2 lambda^2 carries the float rounding of the logarithms (lambda^2 <= 1
    when accepted), and 1 - a0^2 loses ~1.2e-7 absolute to cancellation,
    which the square root divides by 2r."""
    import cmhit_math as cm
    K, n = 6, 4096
    inp = _kp_inputs(12, n, K)
    out = ctx.probe(qcb.F32, qcb.PROBE_KP, inp, 6.0, K)
    skipped = 0
    for row, o in zip(inp, out):
        u = np.ascontiguousarray(row[1:])
        x = np.zeros(4)
        a = qo.lib().qo_kp_injected(6.0, row[0], qo.dp(u), K, qo.dp(x))
        margins, rs = cm.kp_margins(6.0, float(np.float32(row[0])), u, K)
        if min(abs(m) for m in margins) < 1e-5:
            skipped += 1
            continue
        assert a == int(o[0]), (row[0], margins)
        if a >= 0:
            assert abs(o[1] - x[0]) < 4e-6
            assert np.abs(o[2:] - x[1:]).max() < 1e-5 + 4e-7 / max(rs, 1e-12)
    assert skipped <= n // 100


def _cm_inputs(seed, n, K):
    import cmhit_math as cm
    r = np.random.default_rng(seed)
    nslot = 4 * K + 2
    U = [cm.random_su3(r) for _ in range(n)]
    A = [cm.staple_like(r) for _ in range(n)]
    u = r.uniform(size=(n, 3 * nslot))
    for t in range(0, n, 16):                        # some subgroups exhaust their budget
        g = (t // 16) % 3
        u[t, g * nslot:g * nslot + 4 * K:4] = 1.0 - 1e-12
    inp = np.concatenate([np.stack([qo.m3(x) for x in U]), np.stack([qo.m3(x) for x in A]), u], 1)
    return U, A, u, inp


def test_probe_cm_hit(ctx):
    """One Cabibbo-Marinari hit (heatbath_mpi.cpp:11-34) on injected uniforms:
    FP64 against the oracle's cm_hit to 1e-14; FP32 against it within float
    rounding.  The FP32 check excludes float-sensitive items as the KP test
    does (any attempt of any subgroup within 1e-5 of its boundary, at most 2%
    of items) and bounds the rest by 2e-5 + 6e-7 / r_min, r_min the smallest
    sphere radius sqrt(1 - a0^2) of the three draws."""
    import cmhit_math as cm
    K, n = 6, 512
    U, A, u, inp = _cm_inputs(21, n, K)
    out64 = ctx.probe(qcb.F64, qcb.PROBE_CM_HIT, inp, 6.0, K).view(np.complex128).reshape(n, 3, 3)
    out32 = ctx.probe(qcb.F32, qcb.PROBE_CM_HIT, inp, 6.0, K).view(np.complex128).reshape(n, 3, 3)
    skipped = 0
    for t in range(n):
        want = qo.m3(U[t].copy())
        qo.lib().qo_cm_hit_injected(qo.dp(want), qo.dp(qo.m3(A[t])), 6.0, qo.dp(np.ascontiguousarray(u[t])), K)
        want = qo.c3(want)
        assert np.abs(out64[t] - want).max() < 1e-14
        _, info = cm.cm_hit(U[t], A[t], 6.0, u[t], K)
        if min(min(abs(m) for m in margins) for _, _, margins, _ in info) < 1e-5:
            skipped += 1
            continue
        rmin = min([rs for _, x, _, rs in info if x is not None] or [1.0])
        assert np.abs(out32[t] - want).max() < 2e-5 + 6e-7 / max(rmin, 1e-12)
        assert np.abs(out32[t] @ out32[t].conj().T - np.eye(3)).max() < 2e-6
    assert skipped <= n // 50


def test_probe_uniform_words(ctx):
    """The heatbath's uniform of a Philox word is (w + 1/2) 2^-32 exactly in
    double, and the float nearest to it in float (computed without double
    arithmetic on the device)."""
    rng = np.random.default_rng(9)
    edge = [0, 1, 2, 127, 128, 255, 256, 257, 2**23 - 1, 2**23, 2**23 + 1, 2**24 - 1, 2**24, 2**24 + 1,
            2**25 - 1, 2**25, 2**25 + 1, 2**31 - 1, 2**31, 2**32 - 2, 2**32 - 1]
    w = np.concatenate([edge, rng.integers(0, 2**32, size=200000)]).astype(np.float64)
    exact = (w + 0.5) * 2.0 ** -32
    out64 = ctx.probe(qcb.F64, qcb.PROBE_UNIFORM, w[:, None])[:, 0]
    assert np.array_equal(out64, exact)
    out32 = ctx.probe(qcb.F32, qcb.PROBE_UNIFORM, w[:, None])[:, 0]
    assert np.array_equal(out32, exact.astype(np.float32).astype(np.float64))
    assert out32.min() > 0 and out32.max() <= 1.0
    with pytest.raises(qcb.QcbError):
        ctx.probe(qcb.F32, qcb.PROBE_UNIFORM, np.array([[2.0 ** 32]]))


def _uniforms(rng, Vh, n_hits, K):
    return rng.uniform(size=8 * n_hits * 3 * (4 * K + 2) * Vh)


@pytest.mark.parametrize("dims,n_hits", [((4, 4, 4, 4), 1), ((8, 8, 8, 8), 2), ((4, 6, 4, 8), 2)])
def test_heatbath_injected_uniforms(ctx, dims, n_hits):
    """A GPU sweep fed the same uniforms as the oracle reproduces its links to
    1e-12.  Each sweep starts from identical links (the oracle's), so the
    check is one sweep deep: across sweeps, last-bit differences (FMA vs
    separate multiply-add) are amplified ~10^3 per sweep from a random start,
    which a compounded bound below states separately."""
    o = qo.Lattice(dims)
    o.hot_philox(5)
    for s in range(2):                       # leave the random start first
        o.sweep_philox(6.0, 1, 77, s)
    g = qcb.Lattice(ctx, dims)
    K = 4
    rng = np.random.default_rng(3)
    for sweep in range(3):
        g.set_links(o.links)
        u = _uniforms(rng, g.V // 2, n_hits, K)
        ex_g = g.hb_sweep_injected(6.0, n_hits, u, K)
        ex_o = o.sweep_injected(6.0, n_hits, u, K)
        assert ex_g == ex_o
        assert maxrel(g.get_links(), o.links) < REL
    assert abs(g.plaquette() - o.plaquette()) < REL


def test_heatbath_injected_from_hot_start_compounded(ctx):
    """From a Haar-random start the first sweeps are the most sensitive to
    rounding; three compounded sweeps stay within 1e-10."""
    g, o = hot_pair(ctx, (8, 8, 8, 8), 5)
    rng = np.random.default_rng(4)
    for sweep in range(3):
        u = _uniforms(rng, g.V // 2, 2, 4)
        assert g.hb_sweep_injected(6.0, 2, u, 4) == o.sweep_injected(6.0, 2, u, 4)
    assert maxrel(g.get_links(), o.links) < 1e-10
    assert abs(g.plaquette() - o.plaquette()) < 1e-12


@pytest.mark.parametrize("dims", [(4, 4, 6, 4), (10, 6, 4, 8)])
def test_heatbath_philox_matches_oracle(ctx, dims):
    """Device Philox uniforms = the oracle's counter scheme: per-sweep links to 1e-12."""
    g, o = hot_pair(ctx, dims, 9)
    for s in range(2):
        o.sweep_philox(5.9, 2, 123, 100 + s)
    for s in range(3):
        g.set_links(o.links)
        g.hb_sweeps(5.9, 2, 1, seed=123, sweep_idx=10 + s)
        o.sweep_philox(5.9, 2, 123, 10 + s)
        assert maxrel(g.get_links(), o.links) < REL
    # several sweeps in one call follow the same counters
    g.set_links(o.links)
    g.hb_sweeps(5.9, 2, 3, seed=123, sweep_idx=20)
    for s in range(3):
        o.sweep_philox(5.9, 2, 123, 20 + s)
    assert maxrel(g.get_links(), o.links) < 1e-10


def test_heatbath_deterministic_and_su3(ctx):
    a = qcb.Lattice(ctx, (8, 8, 8, 8), qcb.F32)
    b = qcb.Lattice(ctx, (8, 8, 8, 8), qcb.F32)
    for l in (a, b):
        l.hot(4)
        l.hb_sweeps(6.0, 2, 5, seed=99)
    assert np.array_equal(a.get_links(), b.get_links())
    U = a.get_links().reshape(-1, 3, 3)
    assert np.abs(U @ U.conj().transpose(0, 2, 1) - np.eye(3)).max() < 1e-5


def test_fp32_links_stay_su3_without_reunitarize(ctx):
    """FP32 links keep rows 0 and 1 and rebuild row 2 at every load; the update
    kernels re-orthonormalise the two rows before storing, so rounding errors
    do not compound from sweep to sweep (without that, the rebuilt row feeds
    the norm errors back and 300 sweeps overflow)."""
    g = qcb.Lattice(ctx, (8, 8, 8, 8), qcb.F32)
    g.hot(21)
    g.hb_sweeps(6.0, 2, 300, seed=21)
    U = g.get_links().reshape(-1, 3, 3)
    assert np.isfinite(U).all()
    assert np.abs(U @ U.conj().transpose(0, 2, 1) - np.eye(3)).max() < 2e-6
    assert np.abs(np.linalg.det(U) - 1).max() < 2e-6
    g.flow_copy()
    g.flow_rk3(0.02, 150)
    W = g.flow_links().reshape(-1, 3, 3)
    assert np.abs(W @ W.conj().transpose(0, 2, 1) - np.eye(3)).max() < 2e-6


def test_reunitarize_and_translate(ctx):
    dims = (4, 6, 4, 4)
    g, o = hot_pair(ctx, dims, 13)
    rng = np.random.default_rng(4)
    M = o.links + 1e-4 * (rng.normal(size=o.links.shape) + 1j * rng.normal(size=o.links.shape))
    o.set_links(M)
    g.set_links(M)
    g.reunitarize()
    o.reunitarize()
    assert maxrel(g.get_links(), o.links) < 1e-14
    g.set_links(o.links)                      # translation is a pure copy: compare exactly
    g.translate((1, 5, 2, 3))
    o.translate((1, 5, 2, 3))
    assert np.array_equal(g.get_links(), o.links)


# The flow stage's blocks hold whole x rows for Lx/2 = 3 and 5 (30 of the 32
# site lanes) and 16 (all 32): those lattices take the (mu, 0) lower staples
# from LDS; Lx/2 = 7 leaves blocks of 32 sites in partial rows, which form
# every lower staple from the field.  (32, 8, 4, 4) also gives the clover
# its x- and y-planes' shared leaves.
@pytest.mark.parametrize("dims", [(6, 4, 4, 6), (10, 6, 4, 8), (14, 4, 4, 6), (32, 8, 4, 4)])
def test_flow_matches_oracle(ctx, dims):
    g, o = hot_pair(ctx, dims, 17)
    for s in range(2):
        o.sweep_philox(6.0, 1, 8, s)
    g.set_links(o.links)
    tg, fin = g.topo_charge_flowed(0.02, 4, 2, final=True)
    to = o.topo_flowed(0.02, 4, 2)
    assert np.array_equal(tg[:, 0], to[:, 0])
    E = to[:, 2]
    assert np.all(np.abs(tg[:, 2] - E) < REL * E)
    assert np.all(np.abs(tg[:, 1] - to[:, 1]) < REL * E * g.V)
    # the links themselves after the same number of RK3 steps
    g.flow_steps(0.02, 8)
    for _ in range(8):
        o.flow_step(0.02)
    assert maxrel(g.get_links(), o.links) < REL
    assert abs(fin[0] - 0.16) < 1e-15 and abs(fin[2] - o.clover_qe()[1]) < REL * fin[2]
