"""The C-ABI library loads and exports every symbol include/qcb.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import qcb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "qcb.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(qcb_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = qcb.lib()
    names = declared_symbols()
    assert len(names) > 20
    for n in names:
        assert hasattr(lib, n), n


def test_binding_table_matches_header():
    assert sorted(n for n, _, _ in qcb.SIGNATURES) == declared_symbols()


def test_context_creation_reports_missing_gpu():
    """Without a GPU, qcb_ctx_create fails with QCB_EDEVICE (no CPU fallback)."""
    try:
        c = qcb.Context(0)
    except qcb.QcbError as e:
        assert e.code == qcb.EDEVICE
    else:
        c.close()


def test_null_handles_rejected_without_a_device_call():
    """A null handle is QCB_EINVAL before any HIP call, with or without a GPU."""
    lib = qcb.lib()
    out = ctypes.c_double()
    assert lib.qcb_plaquette(None, ctypes.byref(out)) == qcb.EINVAL
    assert lib.qcb_hb_sweeps(None, 6.0, 2, 1, 0, 0) == qcb.EINVAL
    assert lib.qcb_flow_rk3(None, 0.02, 1) == qcb.EINVAL
    assert lib.qcb_lattice_create(None, (ctypes.c_int * 4)(4, 4, 4, 4), qcb.F64, ctypes.byref(ctypes.c_void_p())) \
        == qcb.EINVAL
    assert lib.qcb_ctx_destroy(None) == qcb.EINVAL
