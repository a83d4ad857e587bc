"""C ABI of libasdpocs.so without a GPU: the library loads, exports every
symbol include/asdpocs.h declares, and rejects bad arguments before touching
the device.  The front-end's gateway geometry matches the oracle's bit for
bit, and the MEX-style argument errors carry the reference's messages."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from asd_pocs_mex_cuda_demo_amd import _native as nat
from asd_pocs_mex_cuda_demo_amd import ops

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "asdpocs.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", src)
    return sorted(set(n for n in names if n not in ("if", "sizeof")))


def test_library_exports_every_declared_symbol():
    L = nat.lib()
    decl = declared_functions()
    assert len(decl) >= 20
    for name in decl:
        assert hasattr(L, name), name
    assert sorted(decl) == sorted(nat.EXPORTS)


def test_para_defaults_match_reference():
    p = nat.AsdPara()
    nat.lib().asd_para_defaults(C.byref(p))
    # ASD_POCS_cuda.m:34-96
    assert p.maxnumIter == 5 and p.numOfgradients == 20
    assert p.alpha == 0.02 and p.alpha_red == 0.95 and p.beta == 1.0
    assert p.beta_red == 0.995 and p.r_max == 0.95
    assert p.epsilon == np.finfo(float).eps and p.stopRelChg == np.finfo(float).eps
    assert p.lambda_ == 1.0 and p.precision == 0


def test_bad_arguments_are_rejected_before_the_device():
    L = nat.lib()
    p = nat.AsdPara()
    L.asd_para_defaults(C.byref(p))
    h = C.c_void_p()
    phi = np.ones(4)
    # zero volume
    rc = L.asd_plan_create(C.byref(p), phi.ctypes.data_as(C.c_void_p), 4, C.byref(h))
    assert rc == nat.ASD_ERR_ARG and b"para.nx" in L.asd_last_error()
    p.nx = p.ny = 8
    p.na = 16
    p.SO, p.OD, p.vxlsize, p.detsize = 500.0, 500.0, 1.0, 1.0
    rc = L.asd_plan_create(C.byref(p), phi.ctypes.data_as(C.c_void_p), 0, C.byref(h))
    assert rc == nat.ASD_ERR_ARG and b"phi_vec" in L.asd_last_error()
    p.precision = 7
    rc = L.asd_plan_create(C.byref(p), phi.ctypes.data_as(C.c_void_p), 4, C.byref(h))
    assert rc == nat.ASD_ERR_ARG
    p.precision = 0
    # ny * nz beyond the kernels' 24-bit index factors (the volume itself fits 32 bits)
    p.ny, p.nz = 256, 1 << 17
    rc = L.asd_plan_create(C.byref(p), phi.ctypes.data_as(C.c_void_p), 4, C.byref(h))
    assert rc == nat.ASD_ERR_ARG and b"2^24" in L.asd_last_error()
    p.ny, p.nz = 8, 1
    assert L.asd_tv_min_2d(None, 4, 4, 0.1, 1, 1, None, None) == nat.ASD_ERR_ARG
    assert L.asd_tv_min_3d(None, 4, 4, 4, 0.1, 1, 1, None, None) == nat.ASD_ERR_ARG
    x = np.zeros(4, np.float32)
    assert L.asd_ax_device(x.ctypes.data_as(C.c_void_p), None, None, 1.0, 0.0, None, None, 1.0,
                           1.0, 1.0, 1.0, 0, 1, 1, 1, 1, None) == nat.ASD_ERR_ARG
    # geometry in voxel units: distances, slice and detector pitches positive
    for bad in [dict(SO=0.0), dict(OD=-1.0), dict(dz=0.0), dict(dz=float("nan"))]:
        a = dict(SO=500.0, OD=500.0, dz=1.0)
        a.update(bad)
        assert L.asd_ax_device(x.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p),
                               x.ctypes.data_as(C.c_void_p), 1.0, 0.0,
                               x.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p),
                               a["SO"], a["OD"], 1.0, a["dz"], 2, 2, 1, 2, 1,
                               None) == nat.ASD_ERR_ARG, bad
        assert b"positive" in L.asd_last_error()
    # the void drop-ins report through asd_last_status
    L.Ax_cone_mf_gpu_new(None, None, None, 1.0, 0.0, None, None, 1.0, 1.0, 1.0, 1.0,
                         -1, 1, 1, 1, 1)
    assert L.asd_last_status() == nat.ASD_ERR_ARG
    with pytest.raises(nat.AsdError):
        nat.check_last()
    assert L.asd_plan_iterate(None, 1) == nat.ASD_ERR_ARG
    assert L.asd_plan_set_ax_shape(None, 8, 3) == nat.ASD_ERR_ARG


def test_gateway_geometry_is_bit_identical_to_oracle(orc):
    rng = np.random.default_rng(0)
    for _ in range(50):
        para = dict(nx=64, ny=64, nz=1, na=int(rng.integers(3, 900)), nb=int(rng.integers(1, 9)),
                    SO=float(rng.uniform(100, 900)), OD=float(rng.uniform(100, 900)),
                    vxlsize=float(rng.uniform(0.1, 2)), detsize=float(rng.uniform(0.1, 2)),
                    phi=float(rng.uniform(-7, 7)))
        g = ops.gateway_geometry(para)
        o = orc.gateway_geometry(para["phi"], para["SO"], para["OD"], para["vxlsize"],
                                 para["detsize"], para["na"], para["nb"])
        for k in ("cos_phi", "sin_phi", "SO", "OD", "det", "scale"):
            assert g[k] == o[k], k
        assert np.array_equal(g["y_det"], o["y_det"]) and np.array_equal(g["z_det"], o["z_det"])


@pytest.mark.parametrize("field", ops.GEOM_FIELDS)
def test_missing_para_field_message(field):
    para = dict(nx=4, ny=4, nz=1, SO=1.0, OD=1.0, vxlsize=1.0, phi=0.0, detsize=1.0, na=4, nb=1)
    del para[field]
    with pytest.raises(ops.MexError, match=rf"Error input argument parameter\.{field}!"):
        ops.Ax_cone_gpu(np.zeros(16, np.float32), para)
    with pytest.raises(ops.MexError, match=rf"parameter\.{field}!"):
        ops.Atx_cone_gpu(np.zeros(4, np.float32), para)


def test_size_mismatch_messages():
    para = dict(nx=4, ny=4, nz=1, SO=1.0, OD=1.0, vxlsize=1.0, phi=0.0, detsize=1.0, na=4, nb=1)
    with pytest.raises(ops.MexError, match="Error size of input argument X!"):
        ops.Ax_cone_gpu(np.zeros(15, np.float32), para)
    with pytest.raises(ops.MexError, match="Error size of input argument Y!"):
        ops.Atx_cone_gpu(np.zeros(5, np.float32), para)
    with pytest.raises(ops.MexError, match="double format"):
        ops.imgTVGradMin(np.zeros((4, 4), np.float32), 0.1, 1, 1)
    with pytest.raises(ops.MexError, match="img0 dimensions do not match"):
        ops.imgTVGradMin(np.zeros((4, 4)), 0.1, 1, 1, np.zeros((4, 5)))
    with pytest.raises(ops.MexError, match="phi_vec mismatch"):
        ops.ASD_POCS_cuda(np.zeros((4, 3)), dict(para, phi_vec=[1.0, 2.0]))


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    """No fallback: without the built library every entry point raises
    instead of computing anything another way."""
    monkeypatch.setattr(nat, "_lib", None)
    monkeypatch.setattr(nat, "LIB_PATH", str(tmp_path / "libasdpocs.so"))
    with pytest.raises(RuntimeError, match="is missing"):
        nat.lib()
    with pytest.raises(RuntimeError, match="is missing"):
        ops.Ax_cone_gpu(np.zeros(16, np.float32), dict(nx=4, ny=4, nz=1, SO=1.0, OD=1.0,
                                                        vxlsize=1.0, phi=0.0, detsize=1.0,
                                                        na=4, nb=1))
