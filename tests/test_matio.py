"""The MAT-file (level 5) reader: against scipy.io on files scipy writes
(plain and compressed, every numeric class, logical, 1-D to 3-D, skipped
struct and cell variables), a hand-built big-endian file, malformed input,
and the reference's own test data when the reference checkout is present."""
import io
import os
import struct

import numpy as np
import pytest

from asd_pocs_mex_cuda_demo_amd.matio import MatError, read_mat

REF = "/root/reference/testdata"


def _savemat(d, compress):
    sio = pytest.importorskip("scipy.io")
    buf = io.BytesIO()
    sio.savemat(buf, d, do_compression=compress)
    return buf.getvalue()


@pytest.mark.parametrize("compress", [False, True])
def test_numeric_classes_match_scipy(compress):
    rng = np.random.default_rng(0)
    d = {"d": rng.standard_normal((3, 5)), "s": rng.standard_normal((4, 2, 3)).astype(np.float32),
         "i8": rng.integers(-100, 100, (7,)).astype(np.int8),
         "u16": rng.integers(0, 60000, (2, 9)).astype(np.uint16),
         "i32": rng.integers(-10**9, 10**9, (5, 5)).astype(np.int32),
         "u64": rng.integers(0, 2**62, (3,)).astype(np.uint64),
         "b": rng.random((4, 4)) > 0.5, "scalar": 3.25,
         "st": {"a": 1.0}, "c": np.array([np.arange(3.0), "x"], dtype=object)}
    a = read_mat(_savemat(d, compress))
    import scipy.io as sio
    b = sio.loadmat(io.BytesIO(_savemat(d, compress)), mat_dtype=True)
    for k in ("d", "s", "i8", "u16", "i32", "u64", "b", "scalar"):
        assert a[k].dtype == b[k].dtype, k
        assert a[k].shape == b[k].shape and np.array_equal(a[k], b[k]), k
    assert sorted(a.skipped) == ["c", "st"]


def test_big_endian_file():
    """A file written big-endian ('MI'): one uncompressed 2 x 3 double."""
    vals = np.arange(6, dtype=">f8")
    def el(typ, data):
        pad = (-len(data)) % 8
        return struct.pack(">II", typ, len(data)) + data + b"\0" * pad
    body = el(6, struct.pack(">II", 6, 0)) + el(5, struct.pack(">ii", 2, 3)) + el(1, b"x") + \
        el(9, vals.tobytes())
    hdr = b"MATLAB 5.0 MAT-file".ljust(116, b" ") + b"\0" * 8 + struct.pack(">H", 0x0100) + b"MI"
    a = read_mat(hdr + struct.pack(">II", 14, len(body)) + body)
    assert a["x"].shape == (2, 3) and np.array_equal(a["x"], np.arange(6.0).reshape(2, 3, order="F"))


def test_malformed_input_raises():
    with pytest.raises(MatError):
        read_mat(b"not a mat file")
    good = _savemat({"v": np.ones(40)}, False)
    with pytest.raises(MatError):
        read_mat(good[:-20])


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference test data not present")
def test_reference_testdata():
    """The reader against scipy on the reference's own MAT files, and the
    swine file's geometry scalars against the constants of
    utilities/test_asdpocs.m that the real-geometry GPU tests use (the
    clinical arrays themselves are read here only, never copied)."""
    sio = pytest.importorskip("scipy.io")
    for f in ("swine_2d.mat", "P10240000_head_slice190.mat", "P11570000_abdomen_slice148.mat"):
        a = read_mat(os.path.join(REF, f))
        b = sio.loadmat(os.path.join(REF, f), mat_dtype=True)
        for k, v in a.items():
            assert v.dtype == b[k].dtype and np.array_equal(v, b[k]), (f, k)
    s = read_mat(os.path.join(REF, "swine_2d.mat"))
    from tests.test_gpu_realdata import SWINE
    assert s["prjSwine"].shape == (SWINE["na"], 360)
    assert float(s["sod"][0, 0]) == SWINE["sod"] and float(s["sdd"][0, 0]) == SWINE["sdd"]
    assert float(s["detPixelSize"][0, 0]) == SWINE["dt"] and int(s["na"][0, 0]) == SWINE["na"]
    h = read_mat(os.path.join(REF, "P10240000_head_slice190.mat"))
    assert h["img_slice190"].shape == (512, 512)
