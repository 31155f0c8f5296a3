"""Matrix-algebra statement of one Cabibbo-Marinari heatbath hit, used to pin
the oracle's cm_hit and to screen float-rounding-sensitive cases in the FP32
device tests.  Test infrastructure only.

A hit on U with staple A visits the SU(2) subgroups (0,1), (1,2), (0,2) in
that order (reference src/heatbath/heatbath_mpi.cpp:11-34).  For subgroup
(i,j): W = U A, q = the quaternion part of W's (i,j) block, k = |q|,
V = X(q/k); the Kennedy-Pendleton draw x (src/heatbath/heatbath.cpp:8-36)
gives R = X(mult(x, adj(v))) = V^dag X(x) (the reference's mult composes in
reversed order under its embedding, src/su3/utils.cpp:53-60), embedded in
rows/columns (i,j), and U <- R U.
"""
import numpy as np

import oracle_lib as qo

SUBGROUPS = ((0, 1), (1, 2), (0, 2))


def pauli(q):
    """X(q) = q0 1 + i (q1 s1 + q2 s2 + q3 s3) (src/su3/utils.cpp:33-44)."""
    return np.array([[q[0] + 1j * q[3], q[2] + 1j * q[1]],
                     [-q[2] + 1j * q[1], q[0] - 1j * q[3]]])


def quaternion_part(blk):
    """the q with blk = X(q) + i X(p) for real quaternions q, p"""
    return np.array([(blk[0, 0] + blk[1, 1]).real / 2, (blk[0, 1] + blk[1, 0]).imag / 2,
                     (blk[0, 1] - blk[1, 0]).real / 2, (blk[0, 0] - blk[1, 1]).imag / 2])


def kp_margins(beta, k, u, K):
    """(1 - lambda^2) - delta^2 of each attempt, and the sphere radius sqrt(1 - a0^2)
    of the first accepted one (None if all K reject)."""
    alpha = 2.0 * beta * k / 3.0
    out = []
    for a in range(K):
        d, r1, r2, r3 = u[4 * a:4 * a + 4]
        lam2 = -(np.log(r1) + np.cos(2 * np.pi * r2) ** 2 * np.log(r3)) / (2 * alpha)
        m = (1 - lam2) - d * d
        out.append(m)
        if m >= 0:
            a0 = 1 - 2 * lam2
            return out, np.sqrt(max(0.0, 1 - a0 * a0))
    return out, None


def cm_hit(U, A, beta, u, K):
    """U after one hit with uniforms u[g*(4K+2) + slot]; also the per-subgroup
    (k, x, margins, sphere radius).  x comes from the oracle's KP draw."""
    U = np.array(U, dtype=complex)
    nslot = 4 * K + 2
    info = []
    for g, (i, j) in enumerate(SUBGROUPS):
        W = U @ A
        q = quaternion_part(W[np.ix_([i, j], [i, j])])
        k = float(np.linalg.norm(q))
        ug = np.ascontiguousarray(u[g * nslot:(g + 1) * nslot], dtype=np.float64)
        x = np.zeros(4)
        acc = qo.lib().qo_kp_injected(beta, k, qo.dp(ug), K, qo.dp(x))
        margins, rs = kp_margins(beta, k, ug, K)
        info.append((k, x if acc >= 0 else None, margins, rs))
        if acc < 0:
            continue
        R = np.eye(3, dtype=complex)
        R[np.ix_([i, j], [i, j])] = pauli(q / k).conj().T @ pauli(x)
        U = R @ U
    return U, info


def random_su3(r):
    z = r.normal(size=(3, 3)) + 1j * r.normal(size=(3, 3))
    q, R = np.linalg.qr(z)
    q = q @ np.diag(np.diag(R) / np.abs(np.diag(R)))
    return q / np.linalg.det(q) ** (1 / 3)


def staple_like(r):
    """a sum of six SU(3) matrices, the shape of a staple sum"""
    return sum(random_su3(r) for _ in range(6))
