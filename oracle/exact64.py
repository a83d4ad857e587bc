"""FP64 exact-geometry oracle (test infrastructure only).

Independent of the C restatement: every length here comes from a
Liang-Barsky clip of the infinite line against a closed axis-aligned box,
computed in float64 with numpy.  It is used to pin the FP32 restatement of
the reference projectors (src/AxAtx/Ax_cone_gpu.cu) from the mathematics:

  clip_len            |line ∩ box|, any dimension
  ax2d_exact          exact fan-beam ray sums of a 2-D image (nz = 1, nb = 1)
  ax3d_monotone       the reference's 3-D per-slab rule written as geometry: in each
                      unit slab along the driving axis the chord is replaced by the
                      segment whose transverse and axial extents are both sorted
                      ascending (cu:377-397), then clipped exactly against the cells
  atx3d_exact         exact backprojection over all bins (no footprint)
  tv2d_functional / tv3d_functional
                      the zero-padded isotropic TV functional whose gradient the
                      reference's TV stencils evaluate (imgTVGradMin*.cpp)

Coordinates follow the reference (voxel units): voxel (ix,iy,iz) spans
[ix - nx//2, ix - nx//2 + 1] etc.; the source is (-SO cos, -SO sin, 0) and bin
(ia, ib) is (OD cos - sin*ydet[ia], OD sin + cos*ydet[ia], zdet[ib]).
"""
import numpy as np


def clip_len(p1, p2, lo, hi):
    """Length of the infinite line through p1, p2 inside the closed box [lo, hi].
    Arrays broadcast over leading dims; the last dim is the coordinate."""
    p1 = np.asarray(p1, np.float64)
    p2 = np.asarray(p2, np.float64)
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    d = p2 - p1
    shape = np.broadcast_shapes(p1.shape, p2.shape, lo.shape, hi.shape)[:-1]
    t0 = np.full(shape, -np.inf)
    t1 = np.full(shape, np.inf)
    ok = np.ones(shape, bool)
    for k in range(d.shape[-1]):
        dk = np.broadcast_to(d[..., k], shape)
        pk = np.broadcast_to(p1[..., k], shape)
        lk = np.broadcast_to(lo[..., k], shape)
        hk = np.broadcast_to(hi[..., k], shape)
        par = dk == 0
        ok &= ~par | ((pk >= lk) & (pk <= hk))
        with np.errstate(divide="ignore", invalid="ignore"):
            ta = (lk - pk) / dk
            tb = (hk - pk) / dk
        tmin = np.where(par, -np.inf, np.minimum(ta, tb))
        tmax = np.where(par, np.inf, np.maximum(ta, tb))
        t0 = np.maximum(t0, tmin)
        t1 = np.minimum(t1, tmax)
    L = np.where(ok & (t1 > t0), (t1 - t0), 0.0) * np.linalg.norm(d, axis=-1)
    return L


def fan_rays(phi, SO, OD, ydet, zdet):
    """Source point and bin points (float64) of one view, in voxel units."""
    c, s = np.cos(phi), np.sin(phi)
    src = np.array([-SO * c, -SO * s, 0.0])
    ydet = np.asarray(ydet, np.float64)
    zdet = np.asarray(zdet, np.float64)
    bx = OD * c - s * ydet
    by = OD * s + c * ydet
    B = np.zeros((len(zdet), len(ydet), 3))
    B[:, :, 0] = bx[None, :]
    B[:, :, 1] = by[None, :]
    B[:, :, 2] = zdet[:, None]
    return src, B  # B[ib, ia]


def voxel_boxes(nx, ny, nz):
    ix = np.arange(nx) - nx // 2
    iy = np.arange(ny) - ny // 2
    iz = np.arange(nz) - nz // 2
    Z, Y, X = np.meshgrid(iz, iy, ix, indexing="ij")
    lo = np.stack([X, Y, Z], -1).reshape(-1, 3).astype(np.float64)  # x fastest
    return lo, lo + 1.0


def endpoints(cos_phi, sin_phi, SO, OD, ydet, f32_points=False):
    """Source (x, y) and bin points (bx[ia], by[ia]).  With f32_points the
    points are rounded to float32 exactly as the operators form them
    (cos*(-SO) etc. in FP32, Ax_cone_gpu.cu:259-264), so the ray is the line
    the operators actually intersect; otherwise float64 products."""
    if f32_points:
        c, s = np.float32(cos_phi), np.float32(sin_phi)
        SOf, ODf = np.float32(SO), np.float32(OD)
        yd = np.asarray(ydet, np.float32)
        sx, sy = float(c * (-SOf)), float(s * (-SOf))
        bx = ((c * ODf) - (s * yd)).astype(np.float64)
        by = ((s * ODf) + (c * yd)).astype(np.float64)
        return sx, sy, bx, by
    c, s = float(cos_phi), float(sin_phi)
    ydet = np.asarray(ydet, np.float64)
    return -SO * c, -SO * s, OD * c - s * ydet, OD * s + c * ydet


def system_matrix(cos_phi, sin_phi, SO, OD, ydet, zdet, nx, ny, nz, f32_points=False):
    """Exact lengths A[ray, voxel] (rays ib-major, ia fastest; voxels x fastest),
    using the float32 cos/sin the operators receive, promoted to float64."""
    sx, sy, bx, by = endpoints(cos_phi, sin_phi, SO, OD, ydet, f32_points)
    src = np.array([sx, sy, 0.0])
    zdet = np.asarray(zdet, np.float64)
    rays = []
    for ib in range(len(zdet)):
        for ia in range(len(ydet)):
            rays.append([bx[ia], by[ia], zdet[ib]])
    rays = np.asarray(rays)
    lo, hi = voxel_boxes(nx, ny, nz)
    A = clip_len(src[None, None, :], rays[:, None, :], lo[None], hi[None])
    return A


def system_matrix_2d(cos_phi, sin_phi, SO, OD, ydet, nx, ny, f32_points=False):
    """2-D exact lengths (rays in the z = 0 plane, pixel boxes in x, y)."""
    sx, sy, bx, by = endpoints(cos_phi, sin_phi, SO, OD, ydet, f32_points)
    src = np.array([sx, sy])
    rays = np.stack([bx, by], -1)
    lo3, hi3 = voxel_boxes(nx, ny, 1)
    return clip_len(src[None, None, :], rays[:, None, :], lo3[None, :, :2], hi3[None, :, :2])


def ax3d_monotone(X, cos_phi, sin_phi, SO, OD, ydet, zdet, nx, ny, nz):
    """Ray sums (without the vxl scale) under the reference's 3-D split rule,
    computed as geometry in float64.  X is flat, x fastest.  Returns (y, sr)."""
    c, s = float(cos_phi), float(sin_phi)
    x1, y1 = -SO * c, -SO * s
    X3 = np.asarray(X, np.float64).reshape(nz, ny, nx)
    ys, srs = [], []
    for ib in range(len(zdet)):
        for ia in range(len(ydet)):
            x2 = OD * c - s * float(ydet[ia])
            y2 = OD * s + c * float(ydet[ia])
            z2 = float(zdet[ib])
            xdrv = abs(x1 - x2) > abs(y1 - y2)
            u1, t1 = (x1, y1) if xdrv else (y1, x1)
            du, dt = (x2 - x1, y2 - y1) if xdrv else (y2 - y1, x2 - x1)
            nu, nt = (nx, ny) if xdrv else (ny, nx)
            s1, s2 = dt / du, z2 / du
            L = np.sqrt(1.0 + s1 * s1 + s2 * s2)  # chord length per unit slab
            acc = accl = 0.0
            for iu in range(nu):
                ua = iu - nu // 2
                ta, tb = sorted([t1 + s1 * (ua - u1) + nt // 2, t1 + s1 * (ua + 1 - u1) + nt // 2])
                za, zb = sorted([s2 * (ua - u1) + nz // 2, s2 * (ua + 1 - u1) + nz // 2])
                for ct in range(int(np.floor(ta)), int(np.floor(tb)) + 1):
                    if ct < 0 or ct >= nt:
                        continue
                    for cz in range(int(np.floor(za)), int(np.floor(zb)) + 1):
                        if cz < 0 or cz >= nz:
                            continue
                        # parameter interval of the monotone segment inside the cell
                        lo, hi = 0.0, 1.0
                        if tb > ta:
                            lo = max(lo, (ct - ta) / (tb - ta))
                            hi = min(hi, (ct + 1 - ta) / (tb - ta))
                        if zb > za:
                            lo = max(lo, (cz - za) / (zb - za))
                            hi = min(hi, (cz + 1 - za) / (zb - za))
                        w = max(0.0, hi - lo) * L
                        if w == 0.0:
                            continue
                        ix, iy = (iu, ct) if xdrv else (ct, iu)
                        acc += w * X3[cz, iy, ix]
                        accl += w
            ys.append(acc)
            srs.append(accl)
    return np.asarray(ys), np.asarray(srs)


def tv2d_functional(u, eps=1e-8):
    """Sum over the zero-padded grid r = 1..R+1, c = 1..C+1 of
    sqrt(eps + (U[r,c]-U[r-1,c])^2 + (U[r,c]-U[r,c-1])^2), U = u padded by zeros."""
    R, Cn = u.shape
    U = np.zeros((R + 2, Cn + 2))
    U[1:R + 1, 1:Cn + 1] = u
    P = U[1:, 1:]
    return np.sum(np.sqrt(eps + (P - U[:-1, 1:]) ** 2 + (P - U[1:, :-1]) ** 2))


def tv3d_functional(u, eps=1e-8):
    """3-D analogue with backward differences along x, y, z (zero padding)."""
    Nx, Ny, Nz = u.shape
    U = np.zeros((Nx + 2, Ny + 2, Nz + 2))
    U[1:Nx + 1, 1:Ny + 1, 1:Nz + 1] = u
    P = U[1:, 1:, 1:]
    return np.sum(np.sqrt(eps + (P - U[:-1, 1:, 1:]) ** 2 + (P - U[1:, :-1, 1:]) ** 2
                          + (P - U[1:, 1:, :-1]) ** 2))


def numgrad(f, u, h=1e-6):
    """Central-difference gradient of a scalar function of an array."""
    g = np.zeros_like(u)
    it = np.nditer(u, flags=["multi_index"])
    for _ in it:
        i = it.multi_index
        up = u.copy()
        um = u.copy()
        up[i] += h
        um[i] -= h
        g[i] = (f(up) - f(um)) / (2 * h)
    return g
