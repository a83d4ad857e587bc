"""Inputs simulated on the device (asd_plan_simulate): the phantom equals the
CPU rasteriser (phantoms.py) bit for bit, the noiseless sinogram equals the
forward projector's output view by view and the host encoder's 16-bit
storage, and the Poisson deviates follow the Poisson law (moments and
chi-square over the bins, for the product-of-uniforms branch (mean 4) and
the transformed-rejection branch (means 40 and 1e5)); a seed fixes them."""
import numpy as np
import pytest

from asd_pocs_mex_cuda_demo_amd import ops, phantoms

pytestmark = pytest.mark.gpu


def plan(n=64, nz=1, nv=12, na=96, nb=1, precision=0):
    para = dict(nx=n, ny=n, nz=nz, SO=500.0, OD=400.0, vxlsize=1.0, detsize=1.2, na=na, nb=nb,
                alpha=0.2, numOfgradients=5, maxnumIter=3, precision=precision)
    phi = np.arange(1, nv + 1) * 360.0 / nv
    return para, phi, ops.Plan(para, phi)


@pytest.mark.parametrize("nz", [1, 16])
def test_phantom_and_noiseless_projections(nz):
    rng = np.random.default_rng(4)
    table = phantoms.ellipses(rng)
    para, phi, pl = plan(nz=nz, nb=1 if nz == 1 else 24)
    pl.simulate(table, scale=0.02, I0=0.0)
    ref = (phantoms.shepp_logan_2d(64, table=table) if nz == 1
           else phantoms.shepp_logan_3d(64, table=table, nz=nz)).astype(np.float32)
    assert np.array_equal(pl.reference(), ref.ravel(order="F").astype(np.float64))
    P = pl.projections()
    for i, th in enumerate(phi):
        y, _ = ops.Ax_cone_gpu(ref, dict(para, phi=-th / 180 * np.pi))
        assert np.array_equal(P[:, i], y.astype(np.float64)), i
    pl.close()


def test_u16_storage_matches_host_encoder():
    table = phantoms.ellipses()
    para, phi, p0 = plan()
    p0.simulate(table, I0=0.0)
    P = p0.projections()
    p0.close()
    _, _, p1 = plan(precision=1)
    p1.simulate(table, I0=0.0)
    _, _, p2 = plan(precision=1)
    p2.set_projections(P)
    assert np.array_equal(p1.projections(), p2.projections())
    assert np.abs(p1.projections() - P).max() <= (P.max(0) - P.min(0)).max() / 65535 / 2 * 1.0001
    p1.close()
    p2.close()


def _counts(lam, seed, nv=512, na=1024):
    """Deviates of Poisson(lam) from an empty phantom (every p = 0), clipped
    below at 1 as the log transform does."""
    _, _, pl = plan(nv=nv, na=na)
    pl.simulate(np.zeros((0, 10)), I0=lam, seed=seed)
    v = pl.projections().ravel()
    pl.close()
    return np.rint(lam * np.exp(-v))


@pytest.mark.parametrize("lam", [4.0, 40.0, 1e5])
def test_poisson_law(lam):
    from scipy import stats
    k = _counts(lam, seed=7)
    n = k.size
    # the law of max(K, 1)
    if lam < 1000:
        hi = int(lam + 8 * np.sqrt(lam)) + 1
        kk = np.arange(1, hi)
        pmf = stats.poisson.pmf(kk, lam)
        pmf[0] += stats.poisson.pmf(0, lam)
        obs = np.array([(k == v).sum() for v in kk])
        keep = pmf * n >= 5
        exp_ = pmf[keep] * n
        o = obs[keep]
        chi2 = ((o - exp_) ** 2 / exp_).sum()
        assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-4, chi2
    mean, var = k.mean(), k.var()
    assert abs(mean - lam) <= 5 * np.sqrt(lam / n) + (1.0 if lam < 10 else 0.0), mean
    assert abs(var / lam - 1) <= 5 * np.sqrt(2.0 / n) + (0.05 if lam < 10 else 0.0), var


def test_seed_fixes_the_noise():
    a, b, c = _counts(1e3, 3, nv=8, na=256), _counts(1e3, 3, nv=8, na=256), \
        _counts(1e3, 4, nv=8, na=256)
    assert np.array_equal(a, b) and not np.array_equal(a, c)


def test_reconstruction_on_simulated_data():
    para, phi, pl = plan(n=64, nv=24)
    pl.simulate(phantoms.ellipses(), scale=0.02, I0=1e5, seed=11)
    pl.reset()  # the simulated phantom stays the RMSE reference
    pl.iterate(3)
    _, rmse = pl.out()
    assert len(rmse) == 3 and rmse[-1] < rmse[0] < 0.02
    pl.close()
