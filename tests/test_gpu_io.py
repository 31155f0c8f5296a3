"""ILDG checkpoints staged through the device (reference src/io/ildg.cpp:34-237):
qcb_ildg_save writes exactly the bytes of the host codec (qcb_ildg_write of
the downloaded links), and qcb_ildg_load restores them, in both precisions,
including lattices whose payload spans several staging chunks."""
import ctypes as C

import numpy as np
import pytest

import qcb
from linkstore import assert_fp32_stored

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = qcb.Context(0)
    yield c
    c.close()


def host_write(path, dims, aos):
    d = (C.c_int * 4)(*dims)
    a = np.ascontiguousarray(aos, dtype=np.complex128)
    assert qcb.lib().qcb_ildg_write(str(path).encode(), d, a.ctypes.data_as(C.POINTER(C.c_double))) == 0


def host_read(path, dims):
    V = int(np.prod(dims))
    out = np.zeros((V, 4, 3, 3), dtype=np.complex128)
    d = (C.c_int * 4)(*dims)
    assert qcb.lib().qcb_ildg_read(str(path).encode(), d, out.ctypes.data_as(C.POINTER(C.c_double))) == 0
    return out


# (4,6,4,8): one chunk; (24,24,24,16): 884736 links, two chunks of at most 2^19
@pytest.mark.parametrize("dims", [(4, 6, 4, 8), (24, 24, 24, 16)])
@pytest.mark.parametrize("prec", [qcb.F64, qcb.F32])
def test_staged_save_equals_host_codec(ctx, tmp_path, dims, prec):
    g = qcb.Lattice(ctx, dims, prec)
    g.hot(5)
    g.hb_sweeps(6.0, 1, 1, seed=5)
    g.ildg_save(tmp_path / "dev")
    host_write(tmp_path / "host", dims, g.get_links())
    assert (tmp_path / "dev").read_bytes() == (tmp_path / "host").read_bytes()


@pytest.mark.parametrize("dims", [(4, 6, 4, 8), (24, 24, 24, 16)])
def test_staged_load_restores_links(ctx, tmp_path, dims):
    rng = np.random.default_rng(8)
    V = int(np.prod(dims))
    U = rng.normal(size=(V, 4, 3, 3)) + 1j * rng.normal(size=(V, 4, 3, 3))
    host_write(tmp_path / "cfg", dims, U)
    d = qcb.Lattice(ctx, dims)
    d.ildg_load(tmp_path / "cfg")
    assert np.array_equal(d.get_links(), U)
    f = qcb.Lattice(ctx, dims, qcb.F32)
    f.ildg_load(tmp_path / "cfg")
    assert_fp32_stored(f.get_links(), U)
    # round trip through the staged save
    d.ildg_save(tmp_path / "again")
    assert np.array_equal(host_read(tmp_path / "again", dims), U)


def test_staged_load_rejects_wrong_size(ctx, tmp_path):
    host_write(tmp_path / "small", (4, 4, 4, 4), np.zeros((256, 4, 3, 3), dtype=np.complex128))
    g = qcb.Lattice(ctx, (4, 4, 4, 8))
    with pytest.raises(qcb.QcbError):
        g.ildg_load(tmp_path / "small")
    with pytest.raises(qcb.QcbError):
        g.ildg_load(tmp_path / "missing")
