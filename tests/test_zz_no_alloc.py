"""asd_atx_device_noalloc (the backprojector that allocates nothing; its ray
table lives in a fixed device array) against asd_atx_device, which
allocates its table, through the C ABI on device memory: bit-identical
results on two calls in a row, the second on a smaller view, and both equal
to the CPU oracle's backprojector (src/AxAtx/Ax_cone_gpu.cu:214-331).  A view
beyond the fixed table is rejected before anything is launched."""
import ctypes as C

import numpy as np
import pytest

from asd_pocs_mex_cuda_demo_amd import _native as nat

pytestmark = pytest.mark.gpu


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _vp(t):
    return C.c_void_p(t.data_ptr())


def _call(fn, torch, X, S, Yd, yd, zd, g, dims):
    nx, ny, nz, na, nb = dims
    return fn(_vp(X), _vp(Yd), _vp(S), g["cos_phi"], g["sin_phi"], _vp(yd), _vp(zd), g["SO"],
              g["OD"], g["scale"], g["det"], g["det"], np.float32(1.0), nx, ny, nz, na, nb, None)


def test_noalloc_backprojector_matches_allocating_one(orc):
    torch = pytest.importorskip("torch")
    L = nat.lib()
    rng = np.random.default_rng(21)
    # a cone view, then a smaller fan view (the second call reuses the table)
    for dims, phi, SO, OD, det in [((48, 40, 12, 96, 40), -0.7, 500.0, 500.0, 1.2),
                                   ((32, 32, 1, 45, 1), -2.3, 300.0, 250.0, 1.5)]:
        nx, ny, nz, na, nb = dims
        g = orc.gateway_geometry(phi, SO, OD, 1.0, det, na, nb)
        Y = rng.standard_normal(na * nb).astype(np.float32)
        Yd, yd, zd = _dev(torch, Y), _dev(torch, g["y_det"]), _dev(torch, g["z_det"])
        N = nx * ny * nz
        X1 = torch.full((N,), 7.0, device="cuda")
        S1 = torch.full((N,), 7.0, device="cuda")
        X2 = torch.full((N,), -3.0, device="cuda")
        S2 = torch.full((N,), -3.0, device="cuda")
        torch.cuda.synchronize()
        nat.check(_call(L.asd_atx_device, torch, X1, S1, Yd, yd, zd, g, dims))
        nat.check(_call(L.asd_atx_device_noalloc, torch, X2, S2, Yd, yd, zd, g, dims))
        torch.cuda.synchronize()
        x1, s1, x2, s2 = (t.cpu().numpy() for t in (X1, S1, X2, S2))
        assert np.array_equal(x1, x2) and np.array_equal(s1, s2), dims
        xo, so = orc.atx(Y, g["cos_phi"], g["sin_phi"], g["y_det"], g["z_det"], g["SO"], g["OD"],
                         g["scale"], g["det"], g["det"], np.float32(1.0), nx, ny, nz, na, nb)
        assert np.array_equal(x2, xo) and np.array_equal(s2, so), dims
        assert np.count_nonzero(so) > N // 4


def test_noalloc_rejects_views_beyond_its_table(orc):
    torch = pytest.importorskip("torch")
    L = nat.lib()
    g = orc.gateway_geometry(-0.7, 500.0, 500.0, 1.0, 1.2, 8, 8)
    t = torch.zeros(64, device="cuda")
    dims = (4, 4, 4, 2048, 1024)  # 2^21 bins > 2^20
    rc = _call(L.asd_atx_device_noalloc, torch, t, t, t, t, t, g, dims)
    assert rc == nat.ASD_ERR_ARG
    assert b"1048576" in L.asd_last_error()
