"""ILDG/LIME codec of libqcb (host side, no GPU): byte layout against the
LIME/ILDG specification as the reference writes it (src/io/ildg.cpp:13-136),
read-back, and error paths."""
import ctypes as C
import os
import struct

import numpy as np
import pytest

import qcb


def _write(path, dims, aos):
    d = (C.c_int * 4)(*dims)
    return qcb.lib().qcb_ildg_write(str(path).encode(), d, aos.ctypes.data_as(C.POINTER(C.c_double)))


def _read(path, dims, out):
    d = (C.c_int * 4)(*dims)
    return qcb.lib().qcb_ildg_read(str(path).encode(), d, out.ctypes.data_as(C.POINTER(C.c_double)))


def parse_lime(raw):
    recs, pos = [], 0
    while pos < len(raw):
        magic, ver, flags, n = struct.unpack(">IHHQ", raw[pos:pos + 16])
        typ = raw[pos + 16:pos + 144].split(b"\0")[0].decode()
        recs.append(dict(magic=magic, ver=ver, mb=bool(flags & 0x8000), me=bool(flags & 0x4000), n=n, type=typ,
                         off=pos + 144))
        pos += 144 + n + (-n) % 8
    return recs


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (4, 6, 8, 2)])
def test_layout_and_roundtrip(tmp_path, dims):
    V = int(np.prod(dims))
    rng = np.random.default_rng(1)
    aos = rng.normal(size=V * 72)
    f = tmp_path / "cfg"
    assert _write(f, dims, aos) == 0
    raw = f.read_bytes()
    recs = parse_lime(raw)
    assert [r["type"] for r in recs] == ["ildg-format", "ildg-binary-data"]
    assert all(r["magic"] == 0x456789AB and r["ver"] == 1 for r in recs)
    assert (recs[0]["mb"], recs[0]["me"]) == (True, True)
    assert (recs[1]["mb"], recs[1]["me"]) == (False, True)       # as the reference writes it
    xml = raw[recs[0]["off"]:recs[0]["off"] + recs[0]["n"]].decode()
    assert xml.startswith('<?xml version="1.0" encoding="UTF-8"?>\n<ildgFormat><version>1.0</version>')
    assert "<field>su3gauge</field><precision>64</precision>" in xml
    assert f"<lx>{dims[0]}</lx><ly>{dims[1]}</ly><lz>{dims[2]}</lz><lt>{dims[3]}</lt>" in xml
    assert recs[1]["off"] == 144 + len(xml) + (-len(xml)) % 8 + 144
    assert recs[1]["n"] == V * 72 * 8
    data = np.frombuffer(raw, dtype=">f8", count=V * 72, offset=recs[1]["off"])
    assert np.array_equal(data, aos)
    back = np.zeros(V * 72)
    assert _read(f, dims, back) == 0
    assert np.array_equal(back, aos)


def test_read_errors(tmp_path):
    dims = (4, 4, 4, 4)
    aos = np.ones(256 * 72)
    f = tmp_path / "cfg"
    assert _write(f, dims, aos) == 0
    out = np.zeros(aos.size)
    assert _read(f, (4, 4, 4, 6), np.zeros(384 * 72)) == qcb.EINVAL
    assert b"expected" in qcb.lib().qcb_last_error(None)
    raw = f.read_bytes()
    (tmp_path / "trunc").write_bytes(raw[:-100])
    assert _read(tmp_path / "trunc", dims, out) == qcb.EINVAL
    assert _read(tmp_path / "missing", dims, out) == qcb.EINVAL
    # a file whose only record is not binary data
    (tmp_path / "xmlonly").write_bytes(raw[: 144 + parse_lime(raw)[0]["n"] + (-parse_lime(raw)[0]["n"]) % 8])
    assert _read(tmp_path / "xmlonly", dims, out) == qcb.EINVAL
    assert b"not found" in qcb.lib().qcb_last_error(None)
