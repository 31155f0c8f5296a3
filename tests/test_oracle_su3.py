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


# ---- the reference's hot start, random_su3 (src/su3/utils.cpp:13-30) ----------

RANDOM_SU3_PROBE = r"""
#include <complex>
#include <cstdio>
#include <random>
int main() {
    std::mt19937_64 rng(2024);
    for (int m = 0; m < 64; ++m) {
        std::normal_distribution<double> gauss(0.0, 1.0);   // a fresh object per matrix, as random_su3
        std::complex<double> z[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) z[i][j] = std::complex<double>(gauss(rng), gauss(rng));
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) std::printf("%.17g %.17g\n", z[i][j].real(), z[i][j].imag());
    }
    return 0;
}
"""


def _probe_matrices(compiler, tmp_path):
    import shutil
    import subprocess
    exe_name = os.path.basename(compiler)
    if shutil.which(compiler) is None and not os.path.exists(compiler):
        pytest.skip(f"{compiler} not available")
    src = tmp_path / "rsu3.cpp"
    src.write_text(RANDOM_SU3_PROBE)
    exe = tmp_path / f"rsu3_{exe_name}"
    subprocess.run([compiler, "-O2", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    v = np.array([float(x) for x in out]).reshape(64, 3, 3, 2)
    return v[..., 0] + 1j * v[..., 1]


def eigen_convention_q(A):
    """Q of A = QR in the convention of Eigen's HouseholderQR, built from what
    fixes it mathematically rather than from its tau/essential-vector code:
    step k maps x = a[k:, k] to beta e0 with beta = -sign(Re x0) |x| (sign(0)
    = +) by the unitary that is the identity on the complement of
    w = x - beta e0 (a complex reflection along w: eigenvalue lam on w, 1
    elsewhere, lam fixed by H x = beta e0); Q = H0^dag H1^dag H2^dag.
    Returns Q and the betas (R's real diagonal)."""
    a = np.array(A, dtype=complex)
    tiny = np.finfo(float).tiny
    Q = np.eye(3, dtype=complex)
    betas = []
    for k in range(3):
        x = a[k:, k].copy()
        m = 3 - k
        e0 = np.zeros(m)
        e0[0] = 1.0
        tail = float(np.sum(np.abs(x[1:]) ** 2))
        if tail <= tiny and x[0].imag ** 2 <= tiny:
            H = np.eye(m, dtype=complex)
            beta = x[0].real
        else:
            nrm = np.sqrt(abs(x[0]) ** 2 + tail)
            beta = -nrm if x[0].real >= 0 else nrm
            w = x - beta * e0
            wn = np.linalg.norm(w)
            wh = w / wn
            aa = np.vdot(wh, beta * e0)
            lam = aa / (aa + wn)
            H = np.eye(m, dtype=complex) + (lam - 1) * np.outer(wh, wh.conj())
            assert abs(abs(lam) - 1) < 1e-13                      # unitary
            assert np.abs(H @ x - beta * e0).max() < 1e-13 * nrm  # maps x onto beta e0
        a[k:, :] = H @ a[k:, :]
        Hf = np.eye(3, dtype=complex)
        Hf[k:, k:] = H
        Q = Q @ Hf.conj().T
        betas.append(beta)
    return Q, np.array(betas)


def _principal_cbrt(z):
    return np.abs(z) ** (1.0 / 3.0) * np.exp(1j * np.angle(z) / 3.0)


@pytest.mark.parametrize("imag_first", [False, True])
def test_random_su3_ref_householder_convention(imag_first):
    """qo_random_su3_ref against QR uniqueness: for invertible A, A = QR with R
    upper triangular and a real diagonal of given signs is unique, so Q is the
    positive-diagonal Q of numpy (LAPACK) times diag(sign r_kk), with the signs
    of Eigen's makeHouseholder rule; then Q / det(Q)^(1/3) on the principal
    branch (arg in (-pi/3, pi/3])."""
    m = qo.Mt64(77)
    g = qo.Mt64(77)
    for _ in range(300):
        ws = [g.normal_pair() for _ in range(9)]
        if imag_first:
            A = np.array([b + 1j * a for a, b in ws]).reshape(3, 3)
        else:
            A = np.array([a + 1j * b for a, b in ws]).reshape(3, 3)
        got = m.random_su3(imag_first)
        Q, betas = eigen_convention_q(A)
        q0, r0 = np.linalg.qr(A)
        ph = np.diag(r0) / np.abs(np.diag(r0))
        q0 = q0 @ np.diag(ph)
        assert np.abs(Q - q0 @ np.diag(np.sign(betas))).max() < 1e-13
        R = Q.conj().T @ A
        assert np.abs(np.tril(R, -1)).max() < 1e-13
        assert np.abs(np.diag(R) - betas).max() < 1e-13
        d = np.linalg.det(Q)
        want = Q / _principal_cbrt(d)
        assert np.abs(got - want).max() < 1e-13
        # properties of the output alone
        assert np.abs(got @ got.conj().T - np.eye(3)).max() < 1e-13
        assert abs(np.linalg.det(got) - 1) < 1e-13
        c = _principal_cbrt(d)
        assert -np.pi / 3 < np.angle(c) <= np.pi / 3 + 1e-15
        Rg = got.conj().T @ A * np.conj(c)
        assert np.abs(np.tril(Rg, -1)).max() < 1e-12
        assert np.abs(np.diag(Rg).imag).max() < 1e-12
        assert np.array_equal(np.sign(np.diag(Rg).real), np.sign(betas))


@pytest.mark.parametrize("compiler,imag_first", [("g++", True), ("/opt/rocm/llvm/bin/clang++", False)])
def test_random_su3_ref_draw_order_matches_compiler(compiler, imag_first, tmp_path):
    """Complex(gauss(rng), gauss(rng)) compiled by GCC draws the imaginary part
    first and by clang (the front end of IntelLLVM, the reference cluster's
    compiler) the real part first; the oracle's two orders take exactly the
    probe's matrices."""
    Az = _probe_matrices(compiler, tmp_path)
    m = qo.Mt64(2024)
    other = qo.Mt64(2024)
    for A in Az:
        got = m.random_su3(imag_first)
        Q, _ = eigen_convention_q(A)
        assert np.abs(got - Q / _principal_cbrt(np.linalg.det(Q))).max() < 1e-13
        # the other order gives a different matrix
        assert np.abs(other.random_su3(not imag_first) - got).max() > 1e-6


def test_cm_hit_is_the_subgroup_heatbath_update():
    """qo_cm_hit_injected against the matrix statement of the hit
    (tests/cmhit_math.py): subgroup order, the embedding rows, R = V^dag X(x)
    and U <- R U.  After each subgroup update the quaternion part of the new
    W block is k V^dag X(x) V: its scalar part is k x0, the heatbath's draw of
    Re Tr over the subgroup, and its norm stays k.  A subgroup whose attempts
    all reject is skipped."""
    import cmhit_math as cm
    r = np.random.default_rng(17)
    K = 6
    nslot = 4 * K + 2
    for trial in range(60):
        U = cm.random_su3(r)
        A = cm.staple_like(r)
        u = r.uniform(size=3 * nslot)
        if trial % 4 == 3:                      # force subgroup g to exhaust its budget
            g = trial % 3
            u[g * nslot:g * nslot + 4 * K:4] = 1.0 - 1e-12
        want, info = cm.cm_hit(U, A, 6.0, u, K)
        got = qo.m3(U.copy())   # m3 of a complex128 array is a view of it
        n_exh = qo.lib().qo_cm_hit_injected(qo.dp(got), qo.dp(qo.m3(A)), 6.0, qo.dp(u), K)
        assert n_exh == sum(x is None for _, x, _, _ in info)
        if trial % 4 == 3:
            assert n_exh >= 1
        assert np.abs(qo.c3(got) - want).max() < 1e-13
        # heatbath property of each applied subgroup update
        Ue = np.array(U, dtype=complex)
        for (i, j), (k, x, _, _) in zip(cm.SUBGROUPS, info):
            if x is None:
                continue
            V = cm.pauli(cm.quaternion_part((Ue @ A)[np.ix_([i, j], [i, j])]) / k)
            R = np.eye(3, dtype=complex)
            R[np.ix_([i, j], [i, j])] = V.conj().T @ cm.pauli(x)
            Ue = R @ Ue
            q2 = cm.quaternion_part((Ue @ A)[np.ix_([i, j], [i, j])])
            assert abs(q2[0] - k * x[0]) < 1e-12 * max(1.0, k)
            assert abs(np.linalg.norm(q2) - k) < 1e-12 * max(1.0, k)
