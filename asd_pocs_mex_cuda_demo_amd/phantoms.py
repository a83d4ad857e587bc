"""Synthetic inputs of the benchmark workloads.

Modified Shepp-Logan head phantoms (the published 10-ellipse / 10-ellipsoid
table of Toft's modification of Shepp & Logan), rasterised on the voxel
grid used by the projectors (x fastest), and Poisson-noisy sinograms.

A table row is A, a, b, c, x0, y0, z0, phi, theta, psi (angles in degrees),
the layout of utilities/phantom3d.m:183-200.  The 3-D volume follows
phantom3d.m:76-108: the grid coordinate r_i = (i - (n-1)/2) / ((n-1)/2), a
voxel's point (x, y, z) = (r_j, r_i, r_k) for array index (i, j, k) (MATLAB's
meshgrid puts x along the second index), rotated by the Euler matrix of
(phi, theta, psi) before the offsets are subtracted.  The 2-D image keeps the
2-D phantom's convention (x along the first index, each ellipse rotated by phi
about its own centre; theta and psi unused), which the reference does not
ship (its demos call MATLAB's phantom).
"""
import math

import numpy as np

# A, a, b, c, x0, y0, z0, phi, theta, psi (degrees)
_MSL = np.array([
    [1.0, .6900, .9200, .8100, 0.00, 0.0000, 0.00, 0.0, 0.0, 0.0],
    [-.8, .6624, .8740, .7800, 0.00, -.0184, 0.00, 0.0, 0.0, 0.0],
    [-.2, .1100, .3100, .2200, 0.22, 0.0000, 0.00, -18.0, 0.0, 10.0],
    [-.2, .1600, .4100, .2800, -.22, 0.0000, 0.00, 18.0, 0.0, 10.0],
    [0.1, .2100, .2500, .4100, 0.00, 0.3500, -.15, 0.0, 0.0, 0.0],
    [0.1, .0460, .0460, .0500, 0.00, 0.1000, 0.25, 0.0, 0.0, 0.0],
    [0.1, .0460, .0460, .0500, 0.00, -.1000, 0.25, 0.0, 0.0, 0.0],
    [0.1, .0460, .0230, .0500, -.08, -.6050, 0.00, 0.0, 0.0, 0.0],
    [0.1, .0230, .0230, .0200, 0.00, -.6060, 0.00, 0.0, 0.0, 0.0],
    [0.1, .0230, .0460, .0200, 0.06, -.6050, 0.00, 0.0, 0.0, 0.0],
])
NCOL = 10


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
    for row in e:
        A, a, b, x0, y0, ph = row[0], row[1], row[2], row[4], row[5], row[7]
        t = ph * (math.pi / 180.0)
        ct, st = math.cos(t), math.sin(t)
        xr = (X - x0) * ct + (Y - y0) * st
        yr = -(X - x0) * st + (Y - y0) * ct
        img[(xr / a) ** 2 + (yr / b) ** 2 <= 1.0] += A
    return img * scale


def euler_matrix(phi, theta, psi):
    """The rotation of phantom3d.m:94-101 (angles in degrees), each entry
    formed in the order written there."""
    ph, th, ps = (v * math.pi / 180 for v in (phi, theta, psi))
    cph, sph, cth, sth, cps, sps = (math.cos(ph), math.sin(ph), math.cos(th), math.sin(th),
                                    math.cos(ps), math.sin(ps))
    return np.array([
        [cps * cph - cth * sph * sps, cps * sph + cth * cph * sps, sps * sth],
        [-sps * cph - cth * sph * cps, -sps * sph + cth * cph * cps, cps * sth],
        [sth * sph, -sth * cph, cth]])


def shepp_logan_3d(n, scale=0.02, table=None, nz=None):
    """n x n x nz modified Shepp-Logan volume in phantom3d.m's convention (see
    the module docstring); nz < n takes the central slices (n - nz) // 2 ...
    of the n^3 phantom.  x fastest (array index i = MATLAB's row index)."""
    nz = n if nz is None else nz
    e = ellipses() if table is None else table
    h = (n - 1) / 2
    r = (np.arange(n) - h) / h
    rz = ((np.arange(nz) + (n - nz) // 2) - h) / h
    # array index (i, j): phantom x = r[j], y = r[i]
    Yg, Xg = np.meshgrid(r, r, indexing="ij")
    vol = np.zeros((n, n, nz), np.float64)
    for row in e:
        A, a, b, c, x0, y0, z0 = row[:7]
        m = euler_matrix(row[7], row[8], row[9])
        asq, bsq, csq = a ** 2, b ** 2, c ** 2
        for k in range(nz):
            z = rz[k]
            p1 = m[0, 0] * Xg + m[0, 1] * Yg + m[0, 2] * z
            p2 = m[1, 0] * Xg + m[1, 1] * Yg + m[1, 2] * z
            p3 = m[2, 0] * Xg + m[2, 1] * Yg + m[2, 2] * z
            q = (p1 - x0) ** 2 / asq + (p2 - y0) ** 2 / bsq + (p3 - z0) ** 2 / csq
            vol[:, :, k][q <= 1.0] += A
    return vol * scale


def poisson_log(P, I0, seed):
    """Noisy line integrals: counts ~ Poisson(I0 * exp(-P)) with a seeded
    generator, p = -ln(max(counts, 1) / I0)."""
    rng = np.random.default_rng(seed)
    counts = rng.poisson(I0 * np.exp(-np.asarray(P, np.float64)))
    return -np.log(np.maximum(counts, 1) / I0)
