"""MATLAB-shaped front-end of the ASD-POCS hot path.

Each function has the call shape of the MEX function or MATLAB driver it
replaces, with a dict standing in for the `para` struct:

  Ax_cone_gpu(X, para)           -> (y, sr)   src/AxAtx/Ax_main_test.cpp:19-165
  Atx_cone_gpu(Y, para)          -> (x, sc)   src/AxAtx/Atx_main_test.cpp:20-177
  imgTVGradMin(I, a, n, pos, I0) -> I2        src/tvmin/imgTVGradMin.cpp:16-140
  imgTVGradMin3D(...)            -> I2        src/tvmin/imgTVGradMin3D.cpp:18-189
  ASD_POCS_cuda(P, para)         -> (m, out)  utilities/ASD_POCS_cuda.m:1-182
  SART_cuda_inline(P, para)      -> (m, out)  utilities/SART_cuda_inline.m:1-109

All compute runs in libasdpocs.so on the GPU; argument checks and their
messages follow the reference gateways.
"""
import ctypes as C

import numpy as np

from . import _native as nat

GEOM_FIELDS = ("SO", "OD", "vxlsize", "nx", "ny", "nz", "phi", "detsize", "na", "nb")


class MexError(RuntimeError):
    """Raised where the reference calls mexErrMsgTxt (or prints and returns)."""


def _field(para, name):
    if name not in para or para[name] is None:
        raise MexError(f"Error input argument parameter.{name}!")
    v = para[name]
    return float(np.asarray(v, np.float64).ravel()[0])


def gateway_geometry(para):
    """Voxel-unit geometry exactly as the gateways build it
    (Ax_main_test.cpp:48-161): float casts of every field, cos/sin of the
    float angle, SO/OD/det divided by the voxel size, and the detector tables
    y_det[i] = (i - na/2 + 0.5)*det, z_det[j] = (j - nb/2 + 0.5)*det."""
    vals = {k: _field(para, k) for k in GEOM_FIELDS}
    f32 = np.float32
    SO, OD, scale = f32(vals["SO"]), f32(vals["OD"]), f32(vals["vxlsize"])
    det, phi = f32(vals["detsize"]), f32(vals["phi"])
    nx, ny, nz, na, nb = (int(vals[k]) for k in ("nx", "ny", "nz", "na", "nb"))
    g = dict(nx=nx, ny=ny, nz=nz, na=na, nb=nb)
    g["cos_phi"] = f32(np.cos(np.float64(phi)))
    g["sin_phi"] = f32(np.sin(np.float64(phi)))
    g["SO"] = f32(SO / scale)
    g["OD"] = f32(OD / scale)
    g["det"] = f32(det / scale)
    g["scale"] = scale
    g["dz"] = f32(1.0)
    dd = np.float64(g["det"])
    g["y_det"] = ((np.arange(na, dtype=np.float32).astype(np.float64)
                   - np.float64(f32(na)) / 2.0 + 0.5) * dd).astype(np.float32)
    g["z_det"] = ((np.arange(nb, dtype=np.float32).astype(np.float64)
                   - np.float64(f32(nb)) / 2.0 + 0.5) * dd).astype(np.float32)
    return g


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def Ax_cone_gpu(X, para):
    """[y, sr] = Ax_cone_gpu(single(X), para): one view of the forward projector."""
    g = gateway_geometry(para)
    X = np.ascontiguousarray(np.asarray(X, np.float32).reshape(-1, order="F"))
    if X.size != g["nx"] * g["ny"] * g["nz"]:
        raise MexError("Error size of input argument X!")
    M = g["na"] * g["nb"]
    y = np.zeros(M, np.float32)
    sr = np.zeros(M, np.float32)
    L = nat.lib()
    L.Ax_cone_mf_gpu_new(_ptr(X), _ptr(y), _ptr(sr), g["cos_phi"], g["sin_phi"],
                         _ptr(g["y_det"]), _ptr(g["z_det"]), g["SO"], g["OD"], g["scale"],
                         g["dz"], g["nx"], g["ny"], g["nz"], g["na"], g["nb"])
    nat.check_last()
    return y, sr


def Atx_cone_gpu(Y, para):
    """[x, sc] = Atx_cone_gpu(single(Y), para): one view of the backprojector."""
    g = gateway_geometry(para)
    Y = np.ascontiguousarray(np.asarray(Y, np.float32).reshape(-1, order="F"))
    if Y.size != g["na"] * g["nb"]:
        raise MexError("Error size of input argument Y!")
    N = g["nx"] * g["ny"] * g["nz"]
    x = np.zeros(N, np.float32)
    sc = np.zeros(N, np.float32)
    L = nat.lib()
    L.Atx_cone_mf_gpu_new(_ptr(x), _ptr(Y), _ptr(sc), g["cos_phi"], g["sin_phi"],
                          _ptr(g["y_det"]), _ptr(g["z_det"]), g["SO"], g["OD"], g["scale"],
                          g["det"], g["det"], g["dz"], g["nx"], g["ny"], g["nz"], g["na"],
                          g["nb"])
    nat.check_last()
    return x, sc


def _tv_args(I, I0):
    if not (isinstance(I, np.ndarray) and I.dtype == np.float64):
        raise MexError("Invalid data format. Please make sure the input image data is in double format.")
    if I0 is not None:
        I0 = np.asarray(I0)
        if I0.shape != I.shape:
            raise MexError("img0 dimensions do not match with img")
        I0 = np.ascontiguousarray(I0.astype(np.float64).reshape(-1, order="F"))
    flat = np.ascontiguousarray(I.reshape(-1, order="F"))
    return flat, I0


def imgTVGradMin(I, alpha, nTVIter, pos, I0=None):
    """I2 = imgTVGradMin(I, alpha, nTVIter, pos[, I0]) for a 2-D double image."""
    if I.ndim != 2:
        raise MexError("imgTVGradMin expects a 2-D image")
    flat, i0 = _tv_args(I, I0)
    out = np.zeros_like(flat)
    R, Cn = I.shape
    nat.check(nat.lib().asd_tv_min_2d(_ptr(flat), R, Cn, float(alpha), int(nTVIter), int(pos),
                                      None if i0 is None else _ptr(i0), _ptr(out)))
    return out.reshape(I.shape, order="F")


def imgTVGradMin3D(I, alpha, nTVIter, pos, I0=None):
    """I2 = imgTVGradMin3D(I, alpha, nTVIter, pos[, I0]) for a 3-D double volume."""
    if I.ndim == 2:
        I = I[:, :, None]
    flat, i0 = _tv_args(I, I0 if I0 is None else np.asarray(I0).reshape(I.shape, order="F"))
    out = np.zeros_like(flat)
    Nx, Ny, Nz = I.shape
    nat.check(nat.lib().asd_tv_min_3d(_ptr(flat), Nx, Ny, Nz, float(alpha), int(nTVIter),
                                      int(pos), None if i0 is None else _ptr(i0), _ptr(out)))
    return out.reshape(I.shape, order="F")


def _get(para, key, default):
    v = para.get(key, None)
    if v is None or (hasattr(v, "__len__") and len(v) == 0):
        return default
    return v


def make_para(para):
    """Fill struct asd_para from a para dict with the driver's defaults
    (ASD_POCS_cuda.m:20-96)."""
    p = nat.AsdPara()
    nat.lib().asd_para_defaults(C.byref(p))
    for k in ("nx", "ny", "nz", "na", "nb"):
        setattr(p, k, int(_field(para, k)))
    for k in ("SO", "OD", "vxlsize", "detsize"):
        setattr(p, k, _field(para, k))
    p.maxnumIter = int(_get(para, "maxnumIter", p.maxnumIter))
    p.numOfgradients = int(_get(para, "numOfgradients", p.numOfgradients))
    for k in ("alpha", "alpha_red", "beta", "beta_red", "r_max", "epsilon", "stopRelChg"):
        setattr(p, k, float(_get(para, k, getattr(p, k))))
    # ASD_POCS_cuda.m:24 tests para.initBySART == true, i.e. == 1
    p.initBySART = int(float(_get(para, "initBySART", 0)) == 1.0)
    p.lambda_ = float(_get(para, "lambda", 1.0))
    p.precision = int(_get(para, "precision", 0))
    return p


def _sino(P, para):
    P = np.asarray(P, np.float64)
    if P.ndim == 1:
        P = P[:, None]
    phi = np.ascontiguousarray(np.asarray(para["phi_vec"], np.float64).ravel())
    if P.shape[1] != phi.size:
        raise MexError("Size of Projection and phi_vec mismatch!")
    return np.ascontiguousarray(P.T), phi  # view-major: each view contiguous


def ASD_POCS_cuda(P, para):
    """[m_res, out] = ASD_POCS_cuda(P, para): ASD-POCS on the GPU.

    P is (na*nb, nv) (one column per view).  Returns m_res shaped
    (nx, ny, nz) (the post-SART image of the last iteration) and the `out`
    struct as a dict (IterationNum, dataTolerance, beta, alpha, reconImg,
    RMSE when para['I0'] is given)."""
    Pv, phi = _sino(P, para)
    p = make_para(para)
    N = p.nx * p.ny * p.nz
    m0 = para.get("m0", None)
    if m0 is not None and np.size(m0) > 0:
        m0 = np.ascontiguousarray(np.asarray(m0, np.float64).reshape(-1, order="F"))
    else:
        m0 = None
    I0 = para.get("I0", None)
    if I0 is not None and np.size(I0) > 0:
        I0 = np.ascontiguousarray(np.asarray(I0, np.float64).reshape(-1, order="F"))
    else:
        I0 = None
    m_res = np.zeros(N)
    rmse = np.zeros(max(p.maxnumIter, 1))
    out = nat.AsdOut()
    nat.check(nat.lib().asd_pocs(_ptr(Pv), _ptr(phi), phi.size, C.byref(p),
                                 None if m0 is None else _ptr(m0),
                                 None if I0 is None else _ptr(I0), _ptr(m_res), _ptr(rmse),
                                 C.byref(out)))
    img = m_res.reshape((p.nx, p.ny, p.nz), order="F")
    res = dict(IterationNum=out.IterationNum, dataTolerance=out.dataTolerance, beta=out.beta,
               alpha=out.alpha, reconImg=img)
    if I0 is not None:
        res["RMSE"] = rmse[:out.IterationNum].copy()
    return img, res


def SART_cuda_inline(P, para):
    """[m, out] = SART_cuda_inline(P, para): exactly one SART sweep."""
    Pv, phi = _sino(P, para)
    p = make_para(para)
    N = p.nx * p.ny * p.nz
    m0 = para.get("m0", None)
    m = (np.zeros(N) if m0 is None or np.size(m0) == 0
         else np.ascontiguousarray(np.asarray(m0, np.float64).reshape(-1, order="F")).copy())
    tol = C.c_double(0.0)
    nat.check(nat.lib().asd_sart(_ptr(Pv), _ptr(phi), phi.size, C.byref(p), _ptr(m),
                                 C.byref(tol)))
    out = dict(IterationNum=1, dataTolerance=tol.value, reconImg=m, **{"lambda": p.lambda_})
    return m, out


class Plan:
    """Device-resident ASD-POCS plan (asd_plan_* in include/asdpocs.h)."""

    def __init__(self, para, phi_vec):
        self._L = nat.lib()
        self.para = make_para(para)
        self.phi = np.ascontiguousarray(np.asarray(phi_vec, np.float64).ravel())
        self.N = self.para.nx * self.para.ny * self.para.nz
        self.M = self.para.na * self.para.nb
        h = C.c_void_p()
        nat.check(self._L.asd_plan_create(C.byref(self.para), _ptr(self.phi), self.phi.size,
                                          C.byref(h)))
        self.h = h

    def set_projections(self, P):
        P = np.asarray(P, np.float64)
        Pv = np.ascontiguousarray(P.T if P.shape[0] == self.M else P)
        nat.check(self._L.asd_plan_set_projections(self.h, _ptr(Pv)))

    def reset(self, m0=None, I0=None):
        a = None if m0 is None else np.ascontiguousarray(np.asarray(m0, np.float64).reshape(-1, order="F"))
        b = None if I0 is None else np.ascontiguousarray(np.asarray(I0, np.float64).reshape(-1, order="F"))
        nat.check(self._L.asd_plan_reset(self.h, None if a is None else _ptr(a),
                                         None if b is None else _ptr(b)))

    def iterate(self, n):
        nat.check(self._L.asd_plan_iterate(self.h, int(n)))

    def simulate(self, table, scale=0.02, I0=0.0, seed=0):
        """Phantom (rows A, a, b, c, x0, y0, z0, phi, theta, psi in degrees),
        its projections
        and optional Poisson noise, made on the device (asd_plan_simulate)."""
        t = np.ascontiguousarray(np.asarray(table, np.float64).reshape(-1, 10))
        nat.check(self._L.asd_plan_simulate(self.h, _ptr(t), t.shape[0], float(scale),
                                            float(I0), int(seed)))

    def projections(self):
        """The stored sinogram, decoded, as (na*nb) x nv."""
        P = np.zeros((len(self.phi), self.M))
        nat.check(self._L.asd_plan_get_projections(self.h, _ptr(P)))
        return P.T

    def reference(self):
        """The RMSE reference image (flat, x fastest)."""
        r = np.zeros(self.N)
        nat.check(self._L.asd_plan_get_reference(self.h, _ptr(r)))
        return r

    def sweep(self, lam=1.0):
        tol = C.c_double()
        nat.check(self._L.asd_plan_sweep(self.h, float(lam), C.byref(tol)))
        return tol.value

    def sync(self):
        nat.check(self._L.asd_plan_sync(self.h))

    def image(self):
        m = np.zeros(self.N)
        nat.check(self._L.asd_plan_get_image(self.h, _ptr(m)))
        return m

    def current(self):
        m = np.zeros(self.N)
        nat.check(self._L.asd_plan_get_current(self.h, _ptr(m)))
        return m

    def out(self):
        o = nat.AsdOut()
        nat.check(self._L.asd_plan_get_out(self.h, C.byref(o), None))
        r = np.zeros(max(o.IterationNum, 1))
        nat.check(self._L.asd_plan_get_out(self.h, C.byref(o), _ptr(r)))
        r = r[:o.IterationNum]
        return dict(IterationNum=o.IterationNum, dataTolerance=o.dataTolerance, beta=o.beta,
                    alpha=o.alpha), r

    def set_graph(self, on):
        nat.check(self._L.asd_plan_set_graph(self.h, int(on)))

    def set_timing(self, on):
        nat.check(self._L.asd_plan_set_timing(self.h, int(on)))

    def timing(self, stage):
        ms = C.c_double()
        cnt = C.c_long()
        nat.check(self._L.asd_plan_get_timing(self.h, int(stage), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    def set_atx_variant(self, v):
        nat.check(self._L.asd_plan_set_atx_variant(self.h, int(v)))

    def set_ax_shape(self, lanes_per_ray, rows_per_bundle):
        """Force the forward projector's shape ((0, 0): the default)."""
        nat.check(self._L.asd_plan_set_ax_shape(self.h, int(lanes_per_ray), int(rows_per_bundle)))

    def tuning(self):
        v = C.c_int()
        us = np.zeros(10)
        nat.check(self._L.asd_plan_tuning(self.h, C.byref(v), _ptr(us)))
        names = ["direct", "tile16x16x1", "tile8x8x4", "tile16x4x4", "pairs", "direct2",
                 "direct1", "sorted16", "sorted8x8x4", "sorted8x4x8"]
        return {"atx_variant": names[v.value] if v.value >= 0 else None,
                "atx_us": {n: float(t) for n, t in zip(names, us) if t >= 0}}

    def info(self):
        """(lanes per ray or bundle, detector rows per bundle, hipStream_t)
        of the forward projector."""
        lpr = C.c_int()
        rows = C.c_int()
        st = C.c_void_p()
        nat.check(self._L.asd_plan_info(self.h, C.byref(lpr), C.byref(rows), C.byref(st)))
        return lpr.value, rows.value, st.value

    def close(self):
        if getattr(self, "h", None):
            self._L.asd_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many reconstructions advanced together (asd_batch_* in include/asdpocs.h).
    `paras` is a list of para dicts (each with its own phi_vec)."""

    def __init__(self, paras):
        self._L = nat.lib()
        self.K = len(paras)
        self.paras = paras
        arr = (nat.AsdPara * self.K)()
        self._phis = [np.ascontiguousarray(np.asarray(p["phi_vec"], np.float64).ravel())
                      for p in paras]
        for k, p in enumerate(paras):
            arr[k] = make_para(p)
        self._arr = arr
        ptrs = (C.c_void_p * self.K)(*[ph.ctypes.data for ph in self._phis])
        nvs = (C.c_int * self.K)(*[ph.size for ph in self._phis])
        h = C.c_void_p()
        nat.check(self._L.asd_batch_create(self.K, C.cast(arr, C.c_void_p), C.cast(ptrs, C.c_void_p),
                                           C.cast(nvs, C.c_void_p), C.byref(h)))
        self.h = h

    def set_case(self, c, P, m0=None, I0=None):
        p = self._arr[c]
        P = np.asarray(P, np.float64)
        M = p.na * p.nb
        Pv = np.ascontiguousarray(P.T if P.shape[0] == M else P)
        a = None if m0 is None else np.ascontiguousarray(np.asarray(m0, np.float64).reshape(-1, order="F"))
        b = None if I0 is None else np.ascontiguousarray(np.asarray(I0, np.float64).reshape(-1, order="F"))
        nat.check(self._L.asd_batch_set_case(self.h, c, _ptr(Pv), None if a is None else _ptr(a),
                                             None if b is None else _ptr(b)))

    def iterate(self, n):
        nat.check(self._L.asd_batch_iterate(self.h, int(n)))

    def set_graph(self, on):
        nat.check(self._L.asd_batch_set_graph(self.h, int(on)))

    def tuning(self):
        """{"atx_chunk": footprint pairs per load batch of the backprojector
        stage (None before the first iterate), "atx_us": {1: µs, 2: µs, 4: µs}}
        from the batch's autotuner."""
        ch = C.c_int()
        us = np.zeros(3)
        nat.check(self._L.asd_batch_tuning(self.h, C.byref(ch), _ptr(us)))
        return {"atx_chunk": ch.value or None,
                "atx_us": {k: float(t) for k, t in zip((1, 2, 4), us) if t >= 0}}

    def set_atx_chunk(self, chunk):
        """Forces the backprojector's pairs per load batch (1, 2, 4; 0: tune)."""
        nat.check(self._L.asd_batch_set_atx_chunk(self.h, int(chunk)))

    def sync(self):
        nat.check(self._L.asd_batch_sync(self.h))

    def image(self, c):
        p = self._arr[c]
        m = np.zeros(p.nx * p.ny * p.nz)
        nat.check(self._L.asd_batch_get_image(self.h, c, _ptr(m)))
        return m

    def out(self, c):
        o = nat.AsdOut()
        nat.check(self._L.asd_batch_get_out(self.h, c, C.byref(o), None))
        r = np.zeros(max(o.IterationNum, 1))
        nat.check(self._L.asd_batch_get_out(self.h, c, C.byref(o), _ptr(r)))
        return dict(IterationNum=o.IterationNum, dataTolerance=o.dataTolerance, beta=o.beta,
                    alpha=o.alpha), r[:o.IterationNum]

    def close(self):
        if getattr(self, "h", None):
            self._L.asd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
