"""What an FP32 lattice stores for given links: rows 0 and 1 rounded to
complex64; row 2 is rebuilt on the device as conj(row 0 x row 1), in float."""
import numpy as np


def fp32_stored(U):
    r = np.asarray(U).astype(np.complex64).astype(np.complex128)
    out = r.copy()
    out[..., 2, :] = np.conj(np.cross(r[..., 0, :], r[..., 1, :]))
    return out


def assert_fp32_stored(got, U):
    """rows 0 and 1 bit for bit; row 2 within 1e-6 of the size of the products
    that form each cross-product entry (a few float roundings)"""
    want = fp32_stored(U)
    assert np.array_equal(got[..., :2, :], want[..., :2, :])
    a0, a1 = np.abs(want[..., 0, :]), np.abs(want[..., 1, :])
    size = np.stack([a0[..., (c + 1) % 3] * a1[..., (c + 2) % 3] + a0[..., (c + 2) % 3] * a1[..., (c + 1) % 3]
                     for c in range(3)], axis=-1)
    assert np.all(np.abs(got[..., 2, :] - want[..., 2, :]) <= 1e-6 * size)
