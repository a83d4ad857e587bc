"""GPU parity on the reference's own data and geometries (fixtures written by
tests/golden/make_fixtures.py from its testdata/):

  swine   the measured fan-beam sinogram of testdata/swine_2d.mat with the
          geometry of utilities/test_asdpocs.m (512^2, 1537 bins - an odd
          count, which exercises the backprojector footprint's integer na/2 -
          0.148 mm bins, sod 405.941, sdd 906.251), every 4th of its 360 views
          (the script indexes columns (1:360)*4 of a 360-column array, so the
          views are taken at their own 1-degree angles)
  head    the clinical slice of testdata/P10240000_head_slice190.mat with the
          geometry of demo_head190.m (90 views over 180 degrees, 1024 x 2 bins
          for a one-slice volume: rays off the slice's faces)

Bars as in test_gpu_parity: projectors relative L2 <= 1e-5 (backprojector
bit-identical), reconstructions against the oracle's FP64 driver."""
import os

import numpy as np
import pytest

from asd_pocs_mex_cuda_demo_amd import ops, phantoms

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def swine():
    fx = np.load(os.path.join(GOLD, "swine_2d_90.npz"))
    n, dt, sod, sdd, na = 512, float(fx["det"]), float(fx["sod"]), float(fx["sdd"]), int(fx["na"])
    voxel = sod / sdd * dt * na / n  # test_asdpocs.m: d = sod/sdd*dt*Uy; voxel = d/Nx
    para = dict(nx=n, ny=n, nz=1, SO=sod, OD=sdd - sod, vxlsize=voxel, detsize=dt, na=na, nb=1,
                phi_vec=fx["phi_deg"])
    return para, fx["P"].astype(np.float64)


def head():
    img = np.load(os.path.join(GOLD, "head190.npz"))["img"].astype(np.float64)
    img = img / img.max()  # I3d = img_slice190/max(img_slice190(:))
    n, dt, sod, sdd, nv = 512, 0.148, 405.9410, 906.2510, 90
    Uy, Vz = 2 * n, 2
    voxel = sod / sdd * dt * Uy / n
    para = dict(nx=n, ny=n, nz=1, SO=sod, OD=sdd - sod, vxlsize=voxel, detsize=dt, na=Uy, nb=Vz,
                phi_vec=np.arange(1, nv + 1) * 180.0 / nv)
    return para, img


@pytest.mark.parametrize("case", ["swine", "head"])
def test_projectors_on_reference_geometries(orc, case):
    para, data = swine() if case == "swine" else head()
    rng = np.random.default_rng(11)
    X = (data if case == "head" else rng.random((512, 512))).astype(np.float32)
    for phi_deg in para["phi_vec"][[0, 22, 45, 67]]:
        q = dict(para, phi=-phi_deg / 180 * np.pi)
        y, sr = ops.Ax_cone_gpu(X, q)
        yo, sro = orc.ax_para(X, q)
        assert rel_l2(y, yo) <= 1e-5 and rel_l2(sr, sro) <= 1e-5, (case, phi_deg)
        Y = (yo if case == "head" else rng.standard_normal(para["na"] * para["nb"])).astype(np.float32)
        x, sc = ops.Atx_cone_gpu(Y, q)
        xo, sco = orc.atx_para(Y, q)
        assert np.array_equal(x, xo) and np.array_equal(sc, sco), (case, phi_deg)


def test_swine_reconstruction_matches_oracle(orc):
    """ASD-POCS on the measured swine sinogram (no ground truth: images and
    data residuals are compared)."""
    para, P = swine()
    para.update(maxnumIter=2, alpha=0.2, numOfgradients=5)
    m, out = ops.ASD_POCS_cuda(P, para)
    mo, outo = orc.asd_pocs(P, para)
    assert out["IterationNum"] == outo["IterationNum"] == 2
    assert rel_l2(m.ravel(order="F"), mo) <= 1e-3, rel_l2(m.ravel(order="F"), mo)
    assert abs(out["dataTolerance"] - outo["dataTolerance"]) <= 1e-4 * outo["dataTolerance"]
    assert abs(out["alpha"] - outo["alpha"]) <= 1e-12 * outo["alpha"]


def test_head_reconstruction_matches_oracle(orc):
    """demo_head190.m: projections of the clinical slice, reconstructed with
    the slice as I0; RMSE per iteration against the oracle."""
    para, img = head()
    P = np.zeros((para["na"] * para["nb"], len(para["phi_vec"])))
    for i, th in enumerate(para["phi_vec"]):
        P[:, i] = ops.Ax_cone_gpu(img.astype(np.float32), dict(para, phi=-th / 180 * np.pi))[0]
    para.update(maxnumIter=2, alpha=0.2, alpha_red=0.95, beta_red=1.0, numOfgradients=5, I0=img)
    m, out = ops.ASD_POCS_cuda(P, para)
    mo, outo = orc.asd_pocs(P, para)
    assert np.allclose(out["RMSE"], outo["RMSE"], rtol=1e-4, atol=0), (out["RMSE"], outo["RMSE"])
    assert rel_l2(m.ravel(order="F"), mo) <= 1e-3


def demo3d(n=64):
    """demo_test.m at reduced size: the central 16 slices of an n^3 phantom,
    18 views over 180 degrees (cardinal ones included), Uy = 2 Nx bins by
    Vz = 2 Nz rows of 0.148 mm, sod 405.941, sdd 906.251."""
    vol = phantoms.shepp_logan_3d(n, nz=n)
    vol = vol[:, :, n // 2 - 9:n // 2 + 7]  # I3d(:,:,end/2-8:end/2+7)
    nx, ny, nz = vol.shape
    dt, sod, sdd, nv = 0.148, 405.9410, 906.2510, 18
    Uy, Vz = 2 * nx, 2 * nz
    voxel = sod / sdd * dt * Uy / nx  # d = sod/sdd*dt*Uy; voxel = d/Nx
    para = dict(nx=nx, ny=ny, nz=nz, SO=sod, OD=sdd - sod, vxlsize=voxel, detsize=dt, na=Uy,
                nb=Vz, phi_vec=np.arange(1, nv + 1) * 180.0 / nv)
    return para, vol


def test_demo_test_geometry_projectors_and_reconstruction(orc):
    """demo_test.m's cone geometry: forward projector within 1e-5, the
    backprojector bit-identical (at 90 and 180 degrees too, where rays take
    find_l_3d's a == 0 / b == 0 fallbacks), and two ASD-POCS iterations with
    the phantom as I0 against the oracle."""
    para, vol = demo3d()
    X = np.asfortranarray(vol).ravel(order="F").astype(np.float32)
    rng = np.random.default_rng(5)
    for th in (10.0, 90.0, 130.0, 180.0):
        q = dict(para, phi=-th / 180 * np.pi)
        y, sr = ops.Ax_cone_gpu(X, q)
        yo, sro = orc.ax_para(X, q)
        assert rel_l2(y, yo) <= 1e-5 and rel_l2(sr, sro) <= 1e-5, th
        Y = rng.standard_normal(para["na"] * para["nb"]).astype(np.float32)
        x, sc = ops.Atx_cone_gpu(Y, q)
        xo, sco = orc.atx_para(Y, q)
        assert np.array_equal(x, xo) and np.array_equal(sc, sco), th
    P = np.zeros((para["na"] * para["nb"], len(para["phi_vec"])))
    for i, th in enumerate(para["phi_vec"]):
        P[:, i] = ops.Ax_cone_gpu(X, dict(para, phi=-th / 180 * np.pi))[0]
    para.update(maxnumIter=2, alpha=0.2, alpha_red=0.95, beta_red=1.0, numOfgradients=5, I0=vol)
    m, out = ops.ASD_POCS_cuda(P, para)
    mo, outo = orc.asd_pocs(P, para)
    assert out["IterationNum"] == outo["IterationNum"] == 2
    assert np.allclose(out["RMSE"], outo["RMSE"], rtol=1e-4, atol=0), (out["RMSE"], outo["RMSE"])
    assert rel_l2(m.ravel(order="F"), mo) <= 1e-3, rel_l2(m.ravel(order="F"), mo)
