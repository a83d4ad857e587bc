"""GPU parity on the geometries of the reference's own scripts, with
synthetic data (the reference's measured and clinical arrays are not copied):

  swine   utilities/test_asdpocs.m: a 512^2 fan-beam image, 1537 bins - an
          odd count, which exercises the backprojector footprint's integer
          na/2 - of 0.148 mm, sod 405.941, sdd 906.251 (the geometry of
          testdata/swine_2d.mat; tests/test_matio.py checks these constants
          against the file), views at every 4th of the script's 360
          1-degree angles; the data is a Poisson-noisy (I0 = 1e5)
          modified Shepp-Logan sinogram in place of the measured one
  head    demo_head190.m: 90 views over 180 degrees, 1024 x 2 bins for a
          one-slice 512^2 volume (rays off the slice's faces), the same
          sod / sdd / bin, with the phantom in place of the clinical slice
  demo    demo_test.m at reduced size (the central 16 slices of a 64^3
          phantom, Uy = 2 Nx by Vz = 2 Nz bins)

Bars as in test_gpu_parity: projectors relative L2 <= 1e-5 (backprojector
bit-identical), reconstructions against the oracle's FP64 driver."""
import numpy as np
import pytest

from asd_pocs_mex_cuda_demo_amd import ops, phantoms

pytestmark = pytest.mark.gpu

# utilities/test_asdpocs.m:11-20 (and testdata/swine_2d.mat)
SWINE = dict(n=512, dt=0.148, sod=405.9410, sdd=906.2510, na=1537)


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def swine():
    n, dt, sod, sdd, na = (SWINE[k] for k in ("n", "dt", "sod", "sdd", "na"))
    voxel = sod / sdd * dt * na / n  # test_asdpocs.m: d = sod/sdd*dt*Uy; voxel = d/Nx
    para = dict(nx=n, ny=n, nz=1, SO=sod, OD=sdd - sod, vxlsize=voxel, detsize=dt, na=na, nb=1,
                phi_vec=np.arange(4, 361, 4).astype(np.float64))
    img = phantoms.shepp_logan_2d(n)
    P = np.zeros((na, len(para["phi_vec"])))
    for i, th in enumerate(para["phi_vec"]):
        P[:, i] = ops.Ax_cone_gpu(img.astype(np.float32), dict(para, phi=-th / 180 * np.pi))[0]
    return para, phantoms.poisson_log(P, 1e5, seed=190)


def head():
    n, dt, sod, sdd, nv = 512, 0.148, 405.9410, 906.2510, 90
    Uy, Vz = 2 * n, 2
    voxel = sod / sdd * dt * Uy / n  # demo_head190.m
    para = dict(nx=n, ny=n, nz=1, SO=sod, OD=sdd - sod, vxlsize=voxel, detsize=dt, na=Uy, nb=Vz,
                phi_vec=np.arange(1, nv + 1) * 180.0 / nv)
    return para, phantoms.shepp_logan_2d(n)


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


def test_swine_geometry_reconstruction_matches_oracle(orc):
    """ASD-POCS on a noisy sinogram in the swine geometry (odd bin count):
    images and data residuals against the oracle, and the reference
    arithmetic (precision 4) sweep for sweep."""
    para, P = swine()
    para.update(maxnumIter=2, alpha=0.2, numOfgradients=5)
    m, out = ops.ASD_POCS_cuda(P, para)
    mo, outo = orc.asd_pocs(P, para)
    assert out["IterationNum"] == outo["IterationNum"] == 2
    assert rel_l2(m.ravel(order="F"), mo) <= 1e-3, rel_l2(m.ravel(order="F"), mo)
    assert abs(out["dataTolerance"] - outo["dataTolerance"]) <= 1e-4 * outo["dataTolerance"]
    assert abs(out["alpha"] - outo["alpha"]) <= 1e-12 * outo["alpha"]
    m4, out4 = ops.ASD_POCS_cuda(P, dict(para, precision=4))
    assert rel_l2(m4.ravel(order="F"), mo) <= 1e-12, rel_l2(m4.ravel(order="F"), mo)
    assert out4["dataTolerance"] == pytest.approx(outo["dataTolerance"], rel=1e-13)


def test_head_geometry_reconstruction_matches_oracle(orc):
    """demo_head190.m's geometry: projections of the phantom reconstructed
    with the phantom as I0; RMSE per iteration against the oracle."""
    para, img = head()
    P = np.zeros((para["na"] * para["nb"], len(para["phi_vec"])))
    for i, th in enumerate(para["phi_vec"]):
        P[:, i] = ops.Ax_cone_gpu(img.astype(np.float32), dict(para, phi=-th / 180 * np.pi))[0]
    para.update(maxnumIter=2, alpha=0.2, alpha_red=0.95, beta_red=1.0, numOfgradients=5, I0=img)
    m, out = ops.ASD_POCS_cuda(P, para)
    mo, outo = orc.asd_pocs(P, para)
    assert np.allclose(out["RMSE"], outo["RMSE"], rtol=1e-4, atol=0), (out["RMSE"], outo["RMSE"])
    assert rel_l2(m.ravel(order="F"), mo) <= 1e-3
    m4, out4 = ops.ASD_POCS_cuda(P, dict(para, precision=4))
    assert np.allclose(out4["RMSE"], outo["RMSE"], rtol=1e-12, atol=0)
    assert rel_l2(m4.ravel(order="F"), mo) <= 1e-12


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
    m4, out4 = ops.ASD_POCS_cuda(P, dict(para, precision=4))
    assert np.allclose(out4["RMSE"], outo["RMSE"], rtol=1e-12, atol=0)
    assert rel_l2(m4.ravel(order="F"), mo) <= 1e-12
