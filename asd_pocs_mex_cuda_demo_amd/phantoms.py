"""Synthetic inputs of the benchmark workloads.

Modified Shepp-Logan head phantoms (the published 10-ellipse / 10-ellipsoid
table of Toft's modification of Shepp & Logan), rasterised on the voxel
grid used by the projectors (x fastest), and Poisson-noisy sinograms.
"""
import numpy as np

# A, a, b, c, x0, y0, z0, phi(deg)  (theta = 0, psi as listed below)
_MSL = np.array([
    [1.0, .6900, .9200, .8100, 0.00, 0.0000, 0.00, 0.0],
    [-.8, .6624, .8740, .7800, 0.00, -.0184, 0.00, 0.0],
    [-.2, .1100, .3100, .2200, 0.22, 0.0000, 0.00, -18.0],
    [-.2, .1600, .4100, .2800, -.22, 0.0000, 0.00, 18.0],
    [0.1, .2100, .2500, .4100, 0.00, 0.3500, -.15, 0.0],
    [0.1, .0460, .0460, .0500, 0.00, 0.1000, 0.25, 0.0],
    [0.1, .0460, .0460, .0500, 0.00, -.1000, 0.25, 0.0],
    [0.1, .0460, .0230, .0500, -.08, -.6050, 0.00, 0.0],
    [0.1, .0230, .0230, .0200, 0.00, -.6060, 0.00, 0.0],
    [0.1, .0230, .0460, .0200, 0.06, -.6050, 0.00, 0.0],
])


def ellipses(perturb_rng=None, intensity_jitter=0.10, axis_jitter=0.05):
    """The phantom table, optionally with seeded random intensity (+-10%) and
    axis (+-5%) perturbations per ellipse."""
    e = _MSL.copy()
    if perturb_rng is not None:
        n = len(e)
        e[:, 0] *= 1.0 + perturb_rng.uniform(-intensity_jitter, intensity_jitter, n)
        e[:, 1:4] *= 1.0 + perturb_rng.uniform(-axis_jitter, axis_jitter, (n, 3))
    return e


def shepp_logan_2d(n, scale=0.02, table=None):
    """n x n modified Shepp-Logan image (x = first index), values * scale."""
    e = ellipses() if table is None else table
    c = (np.arange(n) - (n - 1) / 2) / ((n - 1) / 2)
    X, Y = np.meshgrid(c, c, indexing="ij")
    img = np.zeros((n, n))
    for A, a, b, _c, x0, y0, _z0, ph in e:
        t = np.deg2rad(ph)
        xr = (X - x0) * np.cos(t) + (Y - y0) * np.sin(t)
        yr = -(X - x0) * np.sin(t) + (Y - y0) * np.cos(t)
        img[(xr / a) ** 2 + (yr / b) ** 2 <= 1.0] += A
    return img * scale


def shepp_logan_3d(n, scale=0.02, table=None, nz=None):
    """n x n x nz modified Shepp-Logan volume (ellipsoids rotated about z)."""
    nz = n if nz is None else nz
    e = ellipses() if table is None else table
    c = (np.arange(n) - (n - 1) / 2) / ((n - 1) / 2)
    cz = (np.arange(nz) - (nz - 1) / 2) / ((n - 1) / 2)
    X, Y = np.meshgrid(c, c, indexing="ij")
    vol = np.zeros((n, n, nz), np.float64)
    for A, a, b, cc, x0, y0, z0, ph in e:
        t = np.deg2rad(ph)
        xr = (X - x0) * np.cos(t) + (Y - y0) * np.sin(t)
        yr = -(X - x0) * np.sin(t) + (Y - y0) * np.cos(t)
        q = (xr / a) ** 2 + (yr / b) ** 2
        for k in range(nz):
            m = q + ((cz[k] - z0) / cc) ** 2 <= 1.0
            vol[:, :, k][m] += A
    return vol * scale


def poisson_log(P, I0, seed):
    """Noisy line integrals: counts ~ Poisson(I0 * exp(-P)) with a seeded
    generator, p = -ln(max(counts, 1) / I0)."""
    rng = np.random.default_rng(seed)
    counts = rng.poisson(I0 * np.exp(-np.asarray(P, np.float64)))
    return -np.log(np.maximum(counts, 1) / I0)
