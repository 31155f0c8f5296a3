"""Pins of the oracle's event-chain Monte Carlo (reference src/ecmc/ecmc_mpi_cb.cpp,
src/ecmc/ecmc_mpi_cb.h, src/su3/utils.cpp:66-127) by what the mathematics
fixes: the rejection angle solves its defining integral, the two reference
solvers agree, compute_ds and update are the stated matrix expressions, the
near-identity matrices have the embedded-SU(2) product form, and a chain
leaves frozen links and other blocks untouched and keeps links in SU(3)."""
import numpy as np
import pytest
from scipy.integrate import quad

import oracle_lib as qo


def dp(a):
    return qo.dp(np.ascontiguousarray(a))


def c18(m):
    return np.ascontiguousarray(m, dtype=np.complex128).view(np.float64).reshape(18).copy()


def random_su3(rng):
    A = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    Q, R = np.linalg.qr(A)
    Q = Q * (np.diag(R) / np.abs(np.diag(R)))
    return Q / np.linalg.det(Q) ** (1 / 3)


def cumulative_rate(A, B, eps, theta):
    """integral_0^theta max(0, dS/dt) dt with S(t) = A cos(eps t) + B sin(eps t),
    by quadrature between the zeros of dS/dt (located by bisection)"""
    f = lambda t: eps * (-A * np.sin(eps * t) + B * np.cos(eps * t))
    grid = np.linspace(0.0, theta, int(theta / 0.05) + 2)
    edges = [0.0]
    for a, b in zip(grid[:-1], grid[1:]):
        if f(a) * f(b) < 0:
            for _ in range(80):
                m = 0.5 * (a + b)
                a, b = (m, b) if f(a) * f(m) > 0 else (a, m)
            edges.append(0.5 * (a + b))
    edges.append(theta)
    total = 0.0
    for a, b in zip(edges[:-1], edges[1:]):
        if b > a and f(0.5 * (a + b)) > 0:
            total += quad(f, a, b, epsabs=1e-13, epsrel=1e-13)[0]
    return total


def test_reject_angle_solves_its_integral_and_both_solvers_agree():
    rng = np.random.default_rng(1)
    lib = qo.lib()
    for _ in range(300):
        A, B = rng.normal(size=2) * rng.choice([0.05, 1.0, 5.0])
        gamma = rng.exponential() * rng.choice([0.2, 1.0, 20.0])
        eps = int(rng.choice([-1, 1]))
        th = lib.qo_ecmc_reject_fast(A, B, gamma, eps)
        assert th >= 0
        if th < 80:
            assert abs(cumulative_rate(A, B, eps, th) - gamma) < 1e-8 * max(1.0, gamma)
        assert abs(lib.qo_ecmc_reject_exact(A, B, gamma, eps) - th) < 1e-9 * max(1.0, th)


def test_compute_ds_and_update_are_the_stated_products():
    rng = np.random.default_rng(2)
    lib = qo.lib()
    for _ in range(20):
        R, U = random_su3(rng), random_su3(rng)
        P = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        M = R.conj().T @ P @ R
        assert abs(lib.qo_ecmc_compute_ds(dp(c18(P)), dp(c18(R))) - (M[0, 0] - M[1, 1]).imag) < 1e-13
        theta, eps = rng.uniform(0, 7), int(rng.choice([-1, 1]))
        u = c18(U)
        lib.qo_ecmc_update(dp(u), theta, eps, dp(c18(R)))
        xi = eps * theta
        want = R @ np.diag([np.exp(1j * xi), np.exp(-1j * xi), 1.0]) @ R.conj().T @ U
        assert np.abs(u.view(np.complex128).reshape(3, 3) - want).max() < 1e-13


def test_near_identity_matrices():
    rng = np.random.default_rng(3)
    lib = qo.lib()
    sig = [np.array([[0, 1], [1, 0]]), np.array([[0, -1j], [1j, 0]]), np.array([[1, 0], [0, -1]])]
    for eps in (0.15, 0.5):
        r = rng.uniform(-0.5, 0.5, size=9)
        out = np.zeros(18)
        lib.qo_ecmc_random_su3_eps(eps, dp(r), dp(out))
        M = out.view(np.complex128).reshape(3, 3)
        want = np.eye(3, dtype=complex)
        for f, (i, j) in enumerate(((0, 1), (0, 2), (1, 2))):
            q = r[3 * f:3 * f + 3]
            a = eps * q / np.linalg.norm(q)
            # quaternion (a0, a) -> a0 + i a.sigma, the reference's embedding (utils.cpp:33-44)
            X2 = np.sqrt(1 - eps ** 2) * np.eye(2) + 1j * sum(a[k] * sig[k] for k in range(3))
            X = np.eye(3, dtype=complex)
            X[np.ix_([i, j], [i, j])] = X2
            want = want @ X
        assert np.abs(M - want).max() < 1e-14
        assert np.abs(M @ M.conj().T - np.eye(3)).max() < 1e-14


def _frozen_mask(L, b):
    c = np.indices(L[::-1]).reshape(4, -1)[::-1]   # (x, y, z, t) of lex sites
    fr = np.zeros((c.shape[1], 4), dtype=bool)
    for mu in range(4):
        for d in range(4):
            if d != mu:
                fr[:, mu] |= (c[d] % b[d]) == 0
    return fr


@pytest.mark.parametrize("block", [(4, 4, 4, 4), (4, 8, 4, 4)])
def test_philox_chains_touch_only_their_links(block):
    L = (8, 8, 8, 8)
    lat = qo.Lattice(L)
    lat.hot_philox(4)
    p = qo.ecmc_params(block=block)
    ch = lat.ecmc_chains(p)
    U0 = lat.links.copy()
    for s in range(3):
        assert lat.ecmc_philox(p, ch, 11) == len(ch)
    U1 = lat.links
    fr = _frozen_mask(L, block)
    assert np.array_equal(U1[fr], U0[fr])              # frozen links never move
    assert (np.abs(U1[~fr] - U0[~fr]).max(axis=(1, 2)) > 0).mean() > 0.1
    U = U1.reshape(-1, 3, 3)
    assert np.abs(U @ U.conj().transpose(0, 2, 1) - np.eye(3)).max() < 1e-12
    for c in ch:
        assert c.initialized and c.events >= c.lifts > 0
        assert c.set_gen >= 3 and c.rng_n > c.events
    # a chain's block owns its site
    for k, c in enumerate(ch):
        x = np.unravel_index(c.site, L[::-1])[::-1]
        nb = [L[d] // block[d] for d in range(4)]
        bc = (k % nb[0], (k // nb[0]) % nb[1], (k // (nb[0] * nb[1])) % nb[2], k // (nb[0] * nb[1] * nb[2]))
        assert all(x[d] // block[d] == bc[d] for d in range(4))


def test_philox_chains_are_independent_of_each_other():
    """Running the lattice with one chain's block frozen everywhere else gives
    that block the same links: chains do not see each other."""
    L = (8, 8, 8, 8)
    a = qo.Lattice(L)
    a.hot_philox(9)
    b = a.copy()
    p = qo.ecmc_params(block=(4, 4, 4, 4))
    ca, cb = a.ecmc_chains(p), b.ecmc_chains(p)
    a.ecmc_philox(p, ca, 5)
    b.ecmc_philox(p, cb, 5)
    assert np.array_equal(a.links, b.links)     # deterministic
    # perturb a link of block 0's interior: only block 0 may change its result
    c = b.copy()
    cc = c.ecmc_chains(p)
    site = 1 + 8 * (1 + 8 * (1 + 8 * 1))        # (1,1,1,1): block 0, non-frozen for every mu
    c.links[site, 0] = c.links[site, 0] @ np.diag(np.exp(1j * np.array([0.1, -0.05, -0.05])))
    b2 = b.copy()
    b2.ecmc_philox(p, b.ecmc_chains(p), 6)
    c.ecmc_philox(p, cc, 6)
    L_ = np.arange(b.V)
    x = np.stack([L_ % 8, (L_ // 8) % 8, (L_ // 64) % 8, L_ // 512], 1)
    in0 = np.all(x < 4, axis=1)
    assert np.array_equal(b2.links[~in0], c.links[~in0])


def test_reference_mode_chain():
    L = (6, 6, 6, 6)
    lat = qo.Lattice(L)
    lat.hot_philox(2)
    p = qo.ecmc_params(theta_sample=40.0)
    ch = qo.EcmcChain()
    mt = qo.Mt64(12345)
    U0 = lat.links.copy()
    for _ in range(3):
        assert lat.ecmc_ref(p, ch, mt) == 0
    fr = _frozen_mask(L, L)
    assert np.array_equal(lat.links[fr], U0[fr])
    U = lat.links.reshape(-1, 3, 3)
    assert np.abs(U @ U.conj().transpose(0, 2, 1) - np.eye(3)).max() < 1e-12
    c = np.unravel_index(ch.site, L[::-1])[::-1]
    assert ch.events >= ch.lifts > 0
    # the sample event ends exactly on the angle budget
    assert ch.par_refresh_R >= 0 and ch.par_refresh_site >= 0
    assert all(0 <= v < 6 for v in c)
