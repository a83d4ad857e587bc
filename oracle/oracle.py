"""ctypes front-end of the CPU oracle (oracle/asd_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

Functions mirror the reference's call shapes:
  ax / atx          Ax_cone_mf_gpu_new / Atx_cone_mf_gpu_new  (src/AxAtx/Ax_cone_gpu.cu:778, :299)
  ax_para / atx_para  the MEX gateways Ax_cone_gpu / Atx_cone_gpu (src/AxAtx/*_main_test.cpp)
  tv2d / tv3d       imgTVGradMin / imgTVGradMin3D (src/tvmin/imgTVGradMin*.cpp)
  asd_pocs / sart   ASD_POCS_cuda.m / SART_cuda_inline.m
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class OrcPara(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("na", C.c_int),
                ("nb", C.c_int), ("SO", C.c_double), ("OD", C.c_double),
                ("vxlsize", C.c_double), ("detsize", C.c_double),
                ("maxnumIter", C.c_int), ("numOfgradients", C.c_int),
                ("alpha", C.c_double), ("alpha_red", C.c_double), ("beta", C.c_double),
                ("beta_red", C.c_double), ("r_max", C.c_double), ("epsilon", C.c_double),
                ("stopRelChg", C.c_double)]


class OrcOut(C.Structure):
    _fields_ = [("IterationNum", C.c_int), ("dataTolerance", C.c_double),
                ("beta", C.c_double), ("alpha", C.c_double)]


def build():
    """Compile liboracle.so with the Makefile beside this file (with the
    sanitizers when ASDPOCS_SANITIZE=1)."""
    import subprocess
    extra = ["SANITIZE=1"] if os.environ.get("ASDPOCS_SANITIZE") == "1" else []
    subprocess.run(["make", "-B", "-s", "-C", _HERE] + extra, check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        fl, it = C.c_float, C.c_int
        L.orc_find_l.restype = fl
        L.orc_find_l.argtypes = [fl] * 8
        L.orc_find_l_3d.restype = fl
        L.orc_find_l_3d.argtypes = [fl] * 12
        L.orc_gateway_geometry.restype = None
        L.orc_gateway_geometry.argtypes = [C.c_double] * 5 + [it, it] + [C.POINTER(fl)] * 6 + [f32p, f32p]
        L.orc_ax.restype = None
        L.orc_ax.argtypes = [f32p, f32p, f32p, fl, fl, f32p, f32p, fl, fl, fl, fl, it, it, it, it, it]
        L.orc_atx.restype = None
        L.orc_atx.argtypes = [f32p, f32p, f32p, fl, fl, f32p, f32p, fl, fl, fl, fl, fl, fl,
                              it, it, it, it, it]
        lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
        L.orc_ax_rays.restype = None
        L.orc_ax_rays.argtypes = [f32p, f32p, f32p, lp, C.c_long, fl, fl, f32p, f32p, fl, fl, fl,
                                  fl, it, it, it, it, it]
        L.orc_ax_slab_rays.restype = None
        L.orc_ax_slab_rays.argtypes = [f32p, f32p, f32p, lp, C.c_long, fl, fl, f32p, f32p, fl,
                                       fl, fl, fl, it, it, it, it, it, it, it]
        L.orc_atx_voxels.restype = None
        L.orc_atx_voxels.argtypes = [f32p, f32p, f32p, lp, C.c_long, fl, fl, f32p, f32p, fl, fl,
                                     fl, fl, fl, fl, it, it, it, it, it]
        L.orc_sart_view_voxels.restype = None
        L.orc_sart_view_voxels.argtypes = [f64p, f64p] + [C.c_double] * 5 + [it] * 5 + [
            C.c_double, lp, C.c_long, f64p]
        L.orc_tv2d_grad.restype = None
        L.orc_tv2d_grad.argtypes = [f64p, it, it, f64p]
        L.orc_tv3d_grad.restype = None
        L.orc_tv3d_grad.argtypes = [f64p, it, it, it, f64p]
        L.orc_tv2d.restype = it
        L.orc_tv2d.argtypes = [f64p, it, it, C.c_double, it, it, C.c_void_p, f64p]
        L.orc_tv3d.restype = it
        L.orc_tv3d.argtypes = [f64p, it, it, it, C.c_double, it, it, C.c_void_p, f64p]
        L.orc_asd_pocs.restype = it
        L.orc_asd_pocs.argtypes = [f64p, f64p, it, C.POINTER(OrcPara), C.c_void_p, C.c_void_p,
                                   f64p, C.c_void_p, C.POINTER(OrcOut)]
        L.orc_asd_pocs_trace.restype = it
        L.orc_asd_pocs_trace.argtypes = [f64p, f64p, it, C.POINTER(OrcPara), C.c_void_p,
                                         C.c_void_p, f64p, C.c_void_p, C.POINTER(OrcOut), f64p]
        L.orc_sart.restype = C.c_double
        L.orc_sart.argtypes = [f64p, f64p, it, C.POINTER(OrcPara), C.c_double, f64p]
        L.orc_set_sum_order.restype = None
        L.orc_set_sum_order.argtypes = [it]
        L.orc_set_threads.restype = None
        L.orc_set_threads.argtypes = [it]
        L.orc_get_threads.restype = it
        L.orc_get_threads.argtypes = []
        _lib = L
    return _lib


def set_sum_order(order):
    """0: serial FP64 sums of squares (default); 1: pairwise (test-only)."""
    lib().orc_set_sum_order(int(order))


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def find_l(*args):
    return lib().orc_find_l(*[float(a) for a in args])


def find_l_3d(*args):
    return lib().orc_find_l_3d(*[float(a) for a in args])


def gateway_geometry(phi, SO, OD, vxlsize, detsize, na, nb):
    """Voxel-unit geometry exactly as the MEX gateways build it
    (Ax_main_test.cpp:138-161).  Returns a dict of float32 scalars/tables."""
    vals = [C.c_float() for _ in range(6)]
    ydet = np.zeros(na, np.float32)
    zdet = np.zeros(nb, np.float32)
    lib().orc_gateway_geometry(float(phi), float(SO), float(OD), float(vxlsize), float(detsize),
                               int(na), int(nb), *[C.byref(v) for v in vals], ydet, zdet)
    keys = ["cos_phi", "sin_phi", "SO", "OD", "det", "scale"]
    g = {k: np.float32(v.value) for k, v in zip(keys, vals)}
    g["y_det"] = ydet
    g["z_det"] = zdet
    return g


def ax(X, cos_phi, sin_phi, y_det, z_det, SO, OD, scale, dz, nx, ny, nz, na, nb):
    X = np.ascontiguousarray(X, np.float32).ravel(order="F") if X.ndim > 1 else np.ascontiguousarray(X, np.float32)
    y = np.zeros(na * nb, np.float32)
    sr = np.zeros(na * nb, np.float32)
    lib().orc_ax(X, y, sr, cos_phi, sin_phi, np.ascontiguousarray(y_det, np.float32),
                 np.ascontiguousarray(z_det, np.float32), SO, OD, scale, dz, nx, ny, nz, na, nb)
    return y, sr


def atx(Y, cos_phi, sin_phi, y_det, z_det, SO, OD, scale, dy_det, dz_det, dz, nx, ny, nz, na, nb):
    Y = np.ascontiguousarray(Y, np.float32).ravel()
    N = nx * ny * nz
    x = np.zeros(N, np.float32)
    sc = np.zeros(N, np.float32)
    lib().orc_atx(x, Y, sc, cos_phi, sin_phi, np.ascontiguousarray(y_det, np.float32),
                  np.ascontiguousarray(z_det, np.float32), SO, OD, scale, dy_det, dz_det, dz,
                  nx, ny, nz, na, nb)
    return x, sc


def _geom_from_para(para):
    return gateway_geometry(para["phi"], para["SO"], para["OD"], para["vxlsize"],
                            para["detsize"], int(para["na"]), int(para["nb"]))


def ax_para(X, para):
    """[y, sr] = Ax_cone_gpu(single(X), para) restated on the CPU."""
    g = _geom_from_para(para)
    nx, ny, nz, na, nb = (int(para[k]) for k in ("nx", "ny", "nz", "na", "nb"))
    X = np.asarray(X, np.float32).reshape(-1, order="F")
    return ax(X, g["cos_phi"], g["sin_phi"], g["y_det"], g["z_det"], g["SO"], g["OD"],
              g["scale"], np.float32(1.0), nx, ny, nz, na, nb)


def atx_para(Y, para):
    """[x, sc] = Atx_cone_gpu(single(Y), para) restated on the CPU."""
    g = _geom_from_para(para)
    nx, ny, nz, na, nb = (int(para[k]) for k in ("nx", "ny", "nz", "na", "nb"))
    Y = np.asarray(Y, np.float32).reshape(-1, order="F")
    return atx(Y, g["cos_phi"], g["sin_phi"], g["y_det"], g["z_det"], g["SO"], g["OD"],
               g["scale"], g["det"], g["det"], np.float32(1.0), nx, ny, nz, na, nb)


def _dims(para):
    return tuple(int(para[k]) for k in ("nx", "ny", "nz", "na", "nb"))


def ax_rays_para(X, para, ids):
    """Forward-projector sums (y, sr) of the rays `ids` (id = ib*na + ia) of
    the view para['phi'], as ax_para would give them."""
    g = _geom_from_para(para)
    nx, ny, nz, na, nb = _dims(para)
    X = np.ascontiguousarray(np.asarray(X, np.float32).reshape(-1, order="F"))
    ids = np.ascontiguousarray(ids, np.int64)
    y = np.zeros(ids.size, np.float32)
    sr = np.zeros(ids.size, np.float32)
    lib().orc_ax_rays(X, y, sr, ids, ids.size, g["cos_phi"], g["sin_phi"], g["y_det"],
                      g["z_det"], g["SO"], g["OD"], g["scale"], np.float32(1.0), nx, ny, nz, na,
                      nb)
    return y, sr


def ax_slab_rays_para(X_slab, para, ids, z0, nzs):
    """Partial forward-projector sums (y, sr) of the rays `ids` over the
    z-slab [z0, z0 + nzs) of the volume para describes; X_slab holds that
    slab's slices only (x fastest)."""
    g = _geom_from_para(para)
    nx, ny, nz, na, nb = _dims(para)
    X = np.ascontiguousarray(np.asarray(X_slab, np.float32).reshape(-1, order="F"))
    assert X.size == nx * ny * nzs
    ids = np.ascontiguousarray(ids, np.int64)
    y = np.zeros(ids.size, np.float32)
    sr = np.zeros(ids.size, np.float32)
    lib().orc_ax_slab_rays(X, y, sr, ids, ids.size, g["cos_phi"], g["sin_phi"], g["y_det"],
                           g["z_det"], g["SO"], g["OD"], g["scale"], np.float32(1.0), nx, ny, nz,
                           na, nb, int(z0), int(nzs))
    return y, sr


def atx_voxels_para(Y, para, ids):
    """Backprojector sums (x, sc) of the voxels `ids` (MATLAB linear index)
    of the view para['phi'], as atx_para would give them."""
    g = _geom_from_para(para)
    nx, ny, nz, na, nb = _dims(para)
    Y = np.ascontiguousarray(np.asarray(Y, np.float32).reshape(-1, order="F"))
    ids = np.ascontiguousarray(ids, np.int64)
    x = np.zeros(ids.size, np.float32)
    sc = np.zeros(ids.size, np.float32)
    lib().orc_atx_voxels(x, Y, sc, ids, ids.size, g["cos_phi"], g["sin_phi"], g["y_det"],
                         g["z_det"], g["SO"], g["OD"], g["scale"], g["det"], g["det"],
                         np.float32(1.0), nx, ny, nz, na, nb)
    return x, sc


def sart_view_voxels(m, Pv, para, phi_deg, ids, relax=1.0):
    """One SART view update (ASD_POCS_cuda.m:116-131) of the FP64 image m
    with the view's FP64 projection Pv at angle phi_deg (degrees), evaluated
    at the voxels `ids` only."""
    nx, ny, nz, na, nb = _dims(para)
    m = np.ascontiguousarray(np.asarray(m, np.float64).reshape(-1, order="F"))
    Pv = np.ascontiguousarray(np.asarray(Pv, np.float64).ravel())
    ids = np.ascontiguousarray(ids, np.int64)
    out = np.zeros(ids.size)
    lib().orc_sart_view_voxels(m, Pv, -float(phi_deg) / 180 * np.pi, float(para["SO"]),
                               float(para["OD"]), float(para["vxlsize"]), float(para["detsize"]),
                               nx, ny, nz, na, nb, float(relax), ids, ids.size, out)
    return out


def tv2d_grad(u):
    """TV gradient of a (rowN, colN) image in MATLAB column-major order."""
    u = np.asarray(u, np.float64)
    R, Cn = u.shape
    flat = np.ascontiguousarray(u.ravel(order="F"))
    v = np.zeros_like(flat)
    lib().orc_tv2d_grad(flat, R, Cn, v)
    return v.reshape((R, Cn), order="F")


def tv3d_grad(u):
    u = np.asarray(u, np.float64)
    Nx, Ny, Nz = u.shape
    flat = np.ascontiguousarray(u.ravel(order="F"))
    v = np.zeros_like(flat)
    lib().orc_tv3d_grad(flat, Nx, Ny, Nz, v)
    return v.reshape((Nx, Ny, Nz), order="F")


def _opt_ptr(a):
    return None if a is None else a.ctypes.data


def tv2d(I, alpha, nTVIter, pos, I0=None):
    """I2 = imgTVGradMin(I, alpha, nTVIter, pos[, I0]) (column-major images)."""
    I = np.asarray(I, np.float64)
    R, Cn = I.shape
    flat = np.ascontiguousarray(I.ravel(order="F"))
    i0 = None if I0 is None else np.ascontiguousarray(np.asarray(I0, np.float64).ravel(order="F"))
    out = np.zeros_like(flat)
    lib().orc_tv2d(flat, R, Cn, float(alpha), int(nTVIter), int(pos), _opt_ptr(i0), out)
    return out.reshape((R, Cn), order="F")


def tv3d(I, alpha, nTVIter, pos, I0=None):
    I = np.asarray(I, np.float64)
    Nx, Ny, Nz = I.shape
    flat = np.ascontiguousarray(I.ravel(order="F"))
    i0 = None if I0 is None else np.ascontiguousarray(np.asarray(I0, np.float64).ravel(order="F"))
    out = np.zeros_like(flat)
    lib().orc_tv3d(flat, Nx, Ny, Nz, float(alpha), int(nTVIter), int(pos), _opt_ptr(i0), out)
    return out.reshape((Nx, Ny, Nz), order="F")


_DEFAULTS = dict(maxnumIter=5, numOfgradients=20, alpha=0.02, alpha_red=0.95, beta=1.0,
                 beta_red=0.995, r_max=0.95, epsilon=np.finfo(np.float64).eps,
                 stopRelChg=np.finfo(np.float64).eps)


def _orc_para(para):
    p = OrcPara()
    for k in ("nx", "ny", "nz", "na", "nb"):
        setattr(p, k, int(para[k]))
    for k in ("SO", "OD", "vxlsize", "detsize"):
        setattr(p, k, float(para[k]))
    for k, v in _DEFAULTS.items():
        val = para.get(k, None)
        if val is None:
            val = v
        setattr(p, k, int(val) if k in ("maxnumIter", "numOfgradients") else float(val))
    return p


TRACE_COLS = ("dp", "dg", "dd2", "alpha", "beta")


def asd_pocs(P, para, trace=False):
    """[m_res, out] = ASD_POCS_cuda(P, para) restated on the CPU.
    P: (na*nb, nv) array (column = view).  Returns (m_res flat N, out dict).
    trace=True adds out["trace"]: per outer iteration dp, dg, dd2 and the
    alpha and beta left by that iteration's rules (test-only view of the
    step adaptation)."""
    P = np.asarray(P, np.float64)
    nv = P.shape[1]
    phi = np.ascontiguousarray(np.asarray(para["phi_vec"], np.float64).ravel())
    if phi.size != nv:
        raise ValueError("Size of Projection and phi_vec mismatch!")
    Pv = np.ascontiguousarray(P.T)  # each view contiguous
    p = _orc_para(para)
    N = p.nx * p.ny * p.nz
    m0 = para.get("m0", None)
    if m0 is not None:
        m0 = np.ascontiguousarray(np.asarray(m0, np.float64).reshape(-1, order="F"))
    elif para.get("initBySART", None) is not None and float(para["initBySART"]) == 1.0:
        # ASD_POCS_cuda.m:23-25: a non-empty para.initBySART == true (== 1)
        m0 = np.zeros(N)
        lib().orc_sart(Pv, phi, nv, C.byref(p), float(para.get("lambda", 1.0) or 1.0), m0)
    I0 = para.get("I0", None)
    if I0 is not None:
        I0 = np.ascontiguousarray(np.asarray(I0, np.float64).reshape(-1, order="F"))
    m_res = np.zeros(N)
    rmse = np.zeros(max(p.maxnumIter, 1))
    out = OrcOut()
    if trace:
        tr = np.zeros((max(p.maxnumIter, 1), len(TRACE_COLS)))
        lib().orc_asd_pocs_trace(Pv, phi, nv, C.byref(p), _opt_ptr(m0), _opt_ptr(I0), m_res,
                                 rmse.ctypes.data, C.byref(out), tr.ravel())
    else:
        lib().orc_asd_pocs(Pv, phi, nv, C.byref(p), _opt_ptr(m0), _opt_ptr(I0), m_res,
                           rmse.ctypes.data, C.byref(out))
    res = dict(IterationNum=out.IterationNum, dataTolerance=out.dataTolerance, beta=out.beta,
               alpha=out.alpha)
    if trace:
        res["trace"] = {k: tr[:out.IterationNum, j].copy() for j, k in enumerate(TRACE_COLS)}
    if I0 is not None:
        res["RMSE"] = rmse[:out.IterationNum].copy()
    return m_res, res


def sart(P, para):
    """[m, out] = SART_cuda_inline(P, para): one sweep."""
    P = np.asarray(P, np.float64)
    nv = P.shape[1]
    phi = np.ascontiguousarray(np.asarray(para["phi_vec"], np.float64).ravel())
    Pv = np.ascontiguousarray(P.T)
    p = _orc_para(para)
    N = p.nx * p.ny * p.nz
    m0 = para.get("m0", None)
    # a copy: the sweep updates m in place and must not touch the caller's m0
    m = np.zeros(N) if m0 is None else np.array(m0, np.float64).reshape(-1, order="F").copy()
    lam = para.get("lambda", None)
    tol = lib().orc_sart(Pv, phi, nv, C.byref(p), float(1.0 if lam is None else lam), m)
    return m, dict(IterationNum=1, dataTolerance=tol, reconImg=m)
