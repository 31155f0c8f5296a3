"""Pins the oracle's SU(3)/SU(2) algebra against mathematics, not against a
retyped formula: scipy expm fixtures, Pauli-matrix closed forms, group
properties, and a KS test of the Kennedy-Pendleton marginal."""
import os

import numpy as np
import pytest

import oracle_lib as qo

HERE = os.path.dirname(os.path.abspath(__file__))
rng = np.random.default_rng(5)


def haar_su3(r):
    z = r.normal(size=(3, 3)) + 1j * r.normal(size=(3, 3))
    q, R = np.linalg.qr(z)
    q = q @ np.diag(np.diag(R) / np.abs(np.diag(R)))
    return q / np.linalg.det(q) ** (1 / 3)


def test_exp_matches_scipy_expm():
    g = np.load(os.path.join(HERE, "golden", "expm_su3.npz"))
    for Z, c, E in zip(g["Z"], g["coeff"], g["expm"]):
        got = qo.exp_luscher(Z, float(c))
        assert np.abs(got - E).max() < 5e-15 * max(1.0, np.abs(E).max())


def test_exp_is_special_unitary():
    for _ in range(20):
        A = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        Z = qo.proj_lie(A)
        E = qo.exp_luscher(Z, 0.3)
        assert np.allclose(E @ E.conj().T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(E) - 1) < 1e-14


def test_proj_lie_properties():
    A = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    P = qo.proj_lie(A)
    assert np.allclose(P, -P.conj().T, atol=1e-15)
    assert abs(np.trace(P)) < 1e-15
    # idempotent, and kills Hermitian parts and multiples of i*1
    assert np.allclose(qo.proj_lie(P), P, atol=1e-15)
    H = A + A.conj().T
    assert np.allclose(qo.proj_lie(H + 2.5j * np.eye(3)), 0, atol=1e-14)
    # orthogonal projector in the trace inner product: <A - P, X> = 0 for X in su(3)
    X = qo.proj_lie(rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3)))
    assert abs(np.trace((A - P).conj().T @ X).real) < 1e-13


def _embed(q, i, j):
    out = np.zeros(18)
    qo.lib().qo_su2q_to_su3(qo.dp(np.asarray(q, dtype=np.float64)), i, j, qo.dp(out))
    return qo.c3(out)


@pytest.mark.parametrize("ij", [(0, 1), (1, 2), (0, 2)])
def test_su2_embedding_is_pauli_form(ij):
    i, j = ij
    s1 = np.array([[0, 1], [1, 0]]); s2 = np.array([[0, -1j], [1j, 0]]); s3 = np.array([[1, 0], [0, -1]])
    q = rng.normal(size=4)
    block = q[0] * np.eye(2) + 1j * (q[1] * s1 + q[2] * s2 + q[3] * s3)
    X = _embed(q, i, j)
    k = 3 - i - j
    assert X[k, k] == 1
    assert np.allclose(X[np.ix_([i, j], [i, j])], block, atol=0)
    assert X[k, i] == 0 and X[i, k] == 0


def test_quaternion_product_and_roundtrip():
    # The reference's mult (src/su3/utils.cpp:53-60) composes in reversed order
    # under its own embedding: X(mult(q,p)) = X(p) X(q).  su2_step's x*adj(v)
    # is therefore the matrix V^dag X, still an exact heatbath draw because
    # Re Tr(V^dag X k V) = k Re Tr X and the Haar measure is left-invariant.
    lib = qo.lib()
    for _ in range(10):
        q = rng.normal(size=4); p = rng.normal(size=4)
        r = np.zeros(4)
        lib.qo_qmult(qo.dp(q), qo.dp(p), qo.dp(r))
        assert np.allclose(_embed(p, 0, 1) @ _embed(q, 0, 1), _embed(r, 0, 1), atol=1e-14)
        X = _embed(q, 0, 2)
        w = np.array([X[0, 0], X[0, 2], X[2, 0], X[2, 2]]).view(np.float64).copy()
        back = np.zeros(4)
        lib.qo_su2_to_q(qo.dp(w), qo.dp(back))
        assert np.allclose(back, q, atol=1e-15)


def test_kp_sphere_point_and_budget():
    for k in (0.3, 1.0, 2.5):
        u = rng.uniform(size=4 * 8 + 2)
        x = np.zeros(4)
        a = qo.lib().qo_kp_injected(6.0, k, qo.dp(u), 8, qo.dp(x))
        if a >= 0:
            assert abs(np.dot(x, x) - 1) < 1e-14
    # delta = 1 - tiny with r1 small forces rejection of every attempt
    u = np.full(4 * 3 + 2, 0.5); u[0:12:4] = 0.999999; u[1:12:4] = 1e-300
    x = np.full(4, 7.0)
    assert qo.lib().qo_kp_injected(6.0, 1.0, qo.dp(u), 3, qo.dp(x)) == -1
    assert np.all(x == 7.0)


def test_kp_a0_marginal_ks():
    """a0 must follow sqrt(1-a0^2) exp(alpha a0) on [-1,1] (heatbath.cpp:8-36)."""
    from scipy import integrate, stats
    beta, k = 6.0, 0.9
    alpha = 2 * beta * k / 3
    n = 20000
    r = np.random.default_rng(11)
    a0 = []
    lib = qo.lib()
    for _ in range(n):
        u = r.uniform(size=4 * 16 + 2)
        x = np.zeros(4)
        if lib.qo_kp_injected(beta, k, qo.dp(u), 16, qo.dp(x)) >= 0:
            a0.append(x[0])
    dens = lambda t: np.sqrt(max(0.0, 1 - t * t)) * np.exp(alpha * t)
    Z = integrate.quad(dens, -1, 1, limit=200)[0]
    grid = np.linspace(-1, 1, 4001)
    cdf_grid = np.array([0.0] + [integrate.quad(dens, grid[i], grid[i + 1])[0] for i in range(4000)]).cumsum() / Z
    cdf = lambda t: np.interp(t, grid, cdf_grid)
    assert stats.kstest(np.array(a0), cdf).pvalue > 1e-3


def test_reunitarize_link():
    U = haar_su3(rng)
    a = qo.m3(U)
    qo.lib().qo_reunitarize_link(qo.dp(a))
    assert np.allclose(qo.c3(a), U, atol=1e-14)           # fixes SU(3)
    M = U + 1e-3 * (rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3)))
    a = qo.m3(M)
    qo.lib().qo_reunitarize_link(qo.dp(a))
    R = qo.c3(a)
    assert np.allclose(R @ R.conj().T, np.eye(3), atol=1e-14)
    assert abs(np.linalg.det(R) - 1) < 1e-14
    # first two columns are Gram-Schmidt of M's columns: QR with a positive R diagonal
    q, r = np.linalg.qr(M)
    q = q @ np.diag(np.diag(r) / np.abs(np.diag(r)))
    assert np.allclose(R[:, :2], q[:, :2], atol=1e-14)
