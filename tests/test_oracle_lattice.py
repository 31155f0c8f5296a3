"""Pins the oracle's lattice functions by closed forms and invariants:
cold field, constant abelian flux (P, Q, E in closed form), gauge invariance,
the plaquette/staple identity, translation invariance, RK3 order of
convergence, frozen-link bookkeeping and heatbath statistics."""
import numpy as np
import pytest

import oracle_lib as qo

rng = np.random.default_rng(17)


def haar_su3(r, n=1):
    z = r.normal(size=(n, 3, 3)) + 1j * r.normal(size=(n, 3, 3))
    out = np.empty_like(z)
    for k in range(n):
        q, R = np.linalg.qr(z[k])
        q = q @ np.diag(np.diag(R) / np.abs(np.diag(R)))
        out[k] = q / np.linalg.det(q) ** (1 / 3)
    return out


def hot(dims, seed=3):
    l = qo.Lattice(dims)
    l.hot_philox(seed)
    return l


def coords(dims):
    V = int(np.prod(dims))
    s = np.arange(V)
    x = s % dims[0]; y = (s // dims[0]) % dims[1]; z = (s // (dims[0] * dims[1])) % dims[2]
    t = s // (dims[0] * dims[1] * dims[2])
    return np.stack([x, y, z, t], 1)


def lexi(c, dims):
    c = np.asarray(c) % np.asarray(dims)
    return c[..., 0] + dims[0] * (c[..., 1] + dims[1] * (c[..., 2] + dims[2] * c[..., 3]))


def test_cold_field():
    l = qo.Lattice((4, 4, 4, 6))
    assert l.plaquette() == 1.0
    assert l.clover_qe() == (0.0, 0.0)
    assert np.allclose(l.staple(17, 2), 6 * np.eye(3), atol=0)
    l.flow_step(0.02)
    assert np.abs(l.links - np.eye(3)).max() == 0.0


def abelian_field(dims, nf, ng):
    """U diagonal; flux f_k = 2 pi nf_k/(Lx Ly) in plane (0,1) and g_k in (2,3)."""
    Lx, Ly, Lz, Lt = dims
    f = 2 * np.pi * np.asarray(nf) / (Lx * Ly)
    g = 2 * np.pi * np.asarray(ng) / (Lz * Lt)
    c = coords(dims)
    U = np.zeros((len(c), 4, 3, 3), complex)
    for k in range(3):
        U[:, 1, k, k] = np.exp(1j * f[k] * c[:, 0])
        U[:, 0, k, k] = np.where(c[:, 0] == Lx - 1, np.exp(-1j * f[k] * Lx * c[:, 1]), 1.0)
        U[:, 3, k, k] = np.exp(1j * g[k] * c[:, 2])
        U[:, 2, k, k] = np.where(c[:, 2] == Lz - 1, np.exp(-1j * g[k] * Lz * c[:, 3]), 1.0)
    return U, f, g


@pytest.mark.parametrize("dims,nf,ng", [((4, 4, 4, 4), (1, -1, 0), (2, -1, -1)),
                                        ((6, 4, 4, 8), (1, 1, -2), (-1, 0, 1))])
def test_constant_abelian_flux_closed_form(dims, nf, ng):
    U, f, g = abelian_field(dims, nf, ng)
    l = qo.Lattice(dims)
    l.set_links(U)
    V = l.V
    P = (np.cos(f).sum() / 3 + np.cos(g).sum() / 3 + 4) / 6
    assert abs(l.plaquette() - P) < 1e-14
    Q, E = l.clover_qe()
    assert abs(Q - (-V * np.sum(np.sin(f) * np.sin(g)) / (4 * np.pi ** 2))) < 1e-12
    assert abs(E - 0.5 * np.sum(np.sin(f) ** 2 + np.sin(g) ** 2)) < 1e-14


def test_gauge_invariance():
    dims = (4, 4, 4, 4)
    l = hot(dims)
    P0, (Q0, E0) = l.plaquette(), l.clover_qe()
    G = haar_su3(rng, l.V)
    l.gauge_transform(G)
    assert abs(l.plaquette() - P0) < 1e-14
    Q1, E1 = l.clover_qe()
    assert abs(Q1 - Q0) < 1e-12 and abs(E1 - E0) < 1e-12 * max(1, E0)


def test_staple_is_sum_of_plaquettes():
    dims = (4, 4, 6, 4)
    l = hot(dims, 9)
    U = l.links
    c = coords(dims)
    e = np.eye(4, dtype=int)
    tot = 0.0
    for site in (0, 5, 77, l.V - 1):
        for mu in range(4):
            S = l.staple(site, mu)
            want = 0.0
            for nu in range(4):
                if nu == mu:
                    continue
                for base in (c[site], c[site] - e[nu]):
                    b = lexi(base, dims)
                    P = U[b, mu] @ U[lexi(base + e[mu], dims), nu] @ U[lexi(base + e[nu], dims), mu].conj().T @ U[b, nu].conj().T
                    want += np.trace(P).real
            assert abs(np.trace(U[site, mu] @ S).real - want) < 1e-12
    for site in range(l.V):
        for mu in range(4):
            tot += np.trace(U[site, mu] @ l.staple(site, mu)).real
    assert abs(tot - 72 * l.V * l.plaquette()) < 1e-9


def test_translation():
    dims = (4, 6, 4, 4)
    l = hot(dims, 4)
    U0 = l.links.copy()
    P0, (Q0, E0) = l.plaquette(), l.clover_qe()
    l.translate((1, 2, 0, 3))
    c = coords(dims)
    src = lexi(c - np.array([1, 2, 0, 3]), dims)
    assert np.array_equal(l.links, U0[src])
    assert abs(l.plaquette() - P0) < 1e-13
    Q1, E1 = l.clover_qe()
    assert abs(Q1 - Q0) < 1e-12 and abs(E1 - E0) < 1e-12


def _flow_to(l0, eps, T):
    l = l0.copy()
    for _ in range(int(round(T / eps))):
        l.flow_step(eps)
    return l


def test_flow_smooths_and_stays_in_su3():
    l = hot((4, 4, 4, 4), 21)
    p = [l.plaquette()]
    for _ in range(5):
        l.flow_step(0.02)
        p.append(l.plaquette())
    assert all(b > a for a, b in zip(p, p[1:]))
    U = l.links.reshape(-1, 3, 3)
    assert np.abs(U @ U.conj().transpose(0, 2, 1) - np.eye(3)).max() < 1e-13
    assert np.abs(np.linalg.det(U) - 1).max() < 1e-13


def test_flow_rk3_third_order():
    l0 = hot((4, 4, 4, 4), 22)
    T = 0.16
    a = _flow_to(l0, 0.08, T).links.copy()
    b = _flow_to(l0, 0.04, T).links.copy()
    c = _flow_to(l0, 0.02, T).links.copy()
    ratio = np.abs(a - b).max() / np.abs(b - c).max()
    assert 6.0 < ratio < 10.5, ratio   # 2^3 for a third-order integrator


def test_flow_gauge_covariance():
    l = hot((4, 4, 4, 4), 23)
    g = l.copy()
    g.gauge_transform(haar_su3(rng, l.V))
    a = l.topo_flowed(0.02, 4, 2)
    b = g.topo_flowed(0.02, 4, 2)
    assert np.allclose(a, b, rtol=0, atol=1e-11)
    assert np.allclose(a[:, 0], [0, 0.04, 0.08, 0.12])


def test_reference_sweep_frozen_links():
    L = 6
    l = hot((L,) * 4, 8)
    U0 = l.links.copy()
    mt = qo.Mt64(12345)
    visits = l.sweep_ref(6.0, 2, mt)
    assert visits == 4 * L * (L - 1) ** 3
    c = coords((L,) * 4)
    for mu in range(4):
        frozen = np.any((c == 0) & (np.arange(4) != mu), axis=1)
        assert np.array_equal(l.links[frozen, mu], U0[frozen, mu])
        assert not np.any(np.all(l.links[~frozen, mu] == U0[~frozen, mu], axis=(1, 2)))
    U = l.links.reshape(-1, 3, 3)
    assert np.abs(U @ U.conj().transpose(0, 2, 1) - np.eye(3)).max() < 1e-13


def test_injected_sweep_exhaustion_leaves_links():
    dims = (4, 4, 4, 4)
    l = hot(dims, 12)
    U0 = l.links.copy()
    K, n_hits = 2, 1
    u = np.full((8, n_hits, 3, 4 * K + 2, l.V // 2), 0.5)
    u[:, :, :, 0:4 * K:4, :] = 1.0 - 1e-15    # delta
    u[:, :, :, 1:4 * K:4, :] = 1e-300         # r1 -> huge lambda^2 -> reject
    assert l.sweep_injected(6.0, n_hits, np.ascontiguousarray(u.reshape(-1)), K) == 8 * 3 * l.V // 2
    assert np.array_equal(l.links, U0)


def _chain(kind, L, beta, n_therm, n_meas, seed):
    l = qo.Lattice((L,) * 4)
    l.hot_philox(seed)
    mt = qo.Mt64(seed)
    out = []
    for s in range(n_therm + n_meas):
        if kind == "ref":
            l.translate([mt.uniform_int(0, L - 1) for _ in range(4)])
            l.sweep_ref(beta, 2, mt)
        else:
            l.sweep_philox(beta, 2, seed, s)
        if s >= n_therm:
            out.append(l.plaquette())
    return np.array(out)


def _mean_err(x, nbin=10):
    b = x[: len(x) // nbin * nbin].reshape(nbin, -1).mean(1)
    return b.mean(), b.std(ddof=1) / np.sqrt(nbin)


@pytest.mark.slow
def test_checkerboard_and_reference_order_agree():
    """Both update orders sample the same Wilson measure (beta=5.7, 4^4)."""
    a = _chain("ref", 4, 5.7, 40, 200, 1)
    b = _chain("cb", 4, 5.7, 40, 200, 2)
    ma, ea = _mean_err(a)
    mb, eb = _mean_err(b)
    assert abs(ma - mb) < 3.0 * np.hypot(ea, eb) + 1e-4, (ma, ea, mb, eb)
