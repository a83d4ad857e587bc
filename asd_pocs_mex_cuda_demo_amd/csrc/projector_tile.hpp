// Backprojector of the precision variant: tile-local FP64 ray setup.
//
// The reference forms every chord end point relative to the source, about
// SO voxels away, so an FP32 length carries ~1e-4 voxel of cancellation
// error.  Here a workgroup owns a tile of voxels; for every ray that meets
// the tile its line is re-expressed in FP64 relative to the tile centre and
// rounded once to FP32: per voxel only small offsets (|d| < tile size) enter
// the FP32 face tests, and lengths are accurate to ~1e-6.
//
// The face tests and the two-hit rule are the reference's (find_l_3d,
// src/AxAtx/Ax_cone_gpu.cu:141-205) with the same tile-relative quantities:
//   xy = y of the line at x = xc, minus yc  (x faces)   xz = z at x = xc, minus zc
//   yx = x at y = yc, minus xc              (y faces)   yz = z at y = yc, minus zc
//   zx = x at z = zc, minus xc              (z faces)   zy = y at z = zc, minus yc
// A ray whose in-plane line misses the voxel's square has no x/y face hit and
// so no chord at all (a chord through two z faces projects into the square),
// which lets a voxel skip every ib of that ia.  Rays with a == 0 or b == 0
// (the reference's 2-D fallbacks) use the reference's own record instead.
#pragma once

#include "geom_device.hpp"

namespace asd {

// In-plane part of a ray relative to a tile centre (depends on ia only).
struct TileRayXY {
    float e_y, s_ba;   // y of the line at x = x_tc, minus y_tc; dy/dx
    float e_x, s_ab;   // x of the line at y = y_tc, minus x_tc; dx/dy
    int mode;          // RAY_X_FIRST / RAY_Y_FIRST, or a degenerate mode
};

// Axial part relative to the tile centre (depends on ia and ib).
struct TileRayZ {
    float e_zx, s_ca;  // z of the line at x = x_tc, minus z_tc; dz/dx
    float e_zy, s_cb;  // z at y = y_tc, minus z_tc; dz/dy
    float e_xz, s_ac;  // x at z = z_tc, minus x_tc; dx/dz (NaN when c == 0)
    float e_yz, s_bc;  // y at z = z_tc, minus y_tc; dy/dz (NaN when c == 0)
};

// Source and bin point of ray (ydet, zdet) in FP32 as the reference forms
// them (cu:259-264), then everything else in FP64.
__host__ __device__ __forceinline__ void tile_ray_setup(const ViewGeom &g, float ydet, float zdet,
                                                        double xtc, double ytc, double ztc,
                                                        TileRayXY &rxy, TileRayZ &rz)
{
#pragma clang fp contract(off)
    const float x1f = g.cos_phi * (-g.SO), y1f = g.sin_phi * (-g.SO);
    const float x2f = g.cos_phi * g.OD - g.sin_phi * ydet;
    const float y2f = g.sin_phi * g.OD + g.cos_phi * ydet;
    const double x1 = x1f, y1 = y1f, z1 = 0.0;
    const double a = (double)x2f - x1, b = (double)y2f - y1, c = (double)zdet - z1;
    const float fa = x2f - x1f, fb = y2f - y1f;  // the reference's a, b (mode)
    if (fa == 0.0f) rxy.mode = RAY_A0_B0;        // any degenerate: handled by the reference record
    else if (fb == 0.0f) rxy.mode = RAY_B0_C0;
    else rxy.mode = absf(fa) > absf(fb) ? RAY_X_FIRST : RAY_Y_FIRST;
    const double sba = b / a, sab = a / b;
    rxy.s_ba = (float)sba;
    rxy.e_y = (float)(y1 + sba * (xtc - x1) - ytc);
    rxy.s_ab = (float)sab;
    rxy.e_x = (float)(x1 + sab * (ytc - y1) - xtc);
    const double sca = c / a, scb = c / b;
    rz.s_ca = (float)sca;
    rz.e_zx = (float)(z1 + sca * (xtc - x1) - ztc);
    rz.s_cb = (float)scb;
    rz.e_zy = (float)(z1 + scb * (ytc - y1) - ztc);
    if (c != 0.0) {
        const double sac = a / c, sbc = b / c;
        rz.s_ac = (float)sac;
        rz.e_xz = (float)(x1 + sac * (ztc - z1) - xtc);
        rz.s_bc = (float)sbc;
        rz.e_yz = (float)(y1 + sbc * (ztc - z1) - ytc);
    } else {
        const float nan = __builtin_nanf("");
        rz.s_ac = rz.e_xz = rz.s_bc = rz.e_yz = nan;  // no z-face hit
    }
}

// In-plane face conditions of voxel (dx, dy) (offsets from the tile centre):
// bit 0/1: x faces -/+ pass the y test, bit 2/3: y faces -/+ pass the x test.
__host__ __device__ __forceinline__ unsigned tile_inplane(const TileRayXY &r, float dx, float dy,
                                                          float &xy, float &yx)
{
#pragma clang fp contract(off)
    xy = r.e_y + r.s_ba * dx - dy;
    yx = r.e_x + r.s_ab * dy - dx;
    const float xyo = r.s_ba * 0.5f, yxo = r.s_ab * 0.5f;
    unsigned m = 0;
    m |= absf(xy - xyo) <= 0.5f ? 1u : 0u;
    m |= absf(xy + xyo) <= 0.5f ? 2u : 0u;
    m |= absf(yx - yxo) <= 0.5f ? 4u : 0u;
    m |= absf(yx + yxo) <= 0.5f ? 8u : 0u;
    return m;
}

// Chord length of the line inside voxel (dx, dy, dz) given its in-plane
// conditions; faces in the reference's order, first two hits.
__host__ __device__ __forceinline__ float tile_find_l(const TileRayXY &r, const TileRayZ &z,
                                                      unsigned m, float xy, float yx, float dx,
                                                      float dy, float dz, float hz)
{
#pragma clang fp contract(off)
    const float hx = 0.5f, hy = 0.5f;
    const float xz = z.e_zx + z.s_ca * dx - dz, xzo = z.s_ca * hx;
    const float yz = z.e_zy + z.s_cb * dy - dz, yzo = z.s_cb * hy;
    const float zx = z.e_xz + z.s_ac * dz - dx, zxo = z.s_ac * hz;
    const float zy = z.e_yz + z.s_bc * dz - dy, zyo = z.s_bc * hz;
    const float xyo = r.s_ba * hx, yxo = r.s_ab * hy;
    const bool f[6] = {(m & 1u) && absf(xz - xzo) <= hz,
                       (m & 2u) && absf(xz + xzo) <= hz,
                       (m & 4u) && absf(yz - yzo) <= hz,
                       (m & 8u) && absf(yz + yzo) <= hz,
                       absf(zx - zxo) <= hx && absf(zy - zyo) <= hy,
                       absf(zx + zxo) <= hx && absf(zy + zyo) <= hy};
    const float cx[6] = {-hx, hx, yx - yxo, yx + yxo, zx - zxo, zx + zxo};
    const float cy[6] = {xy - xyo, xy + xyo, -hy, hy, zy - zyo, zy + zyo};
    const float cz[6] = {xz - xzo, xz + xzo, yz - yzo, yz + yzo, -hz, hz};
    float ex, ey, ez;
    const bool two = r.mode == RAY_X_FIRST ? chord6<true>(f, cx, cy, cz, ex, ey, ez)
                                           : chord6<false>(f, cx, cy, cz, ex, ey, ez);
    if (!two) return 0.0f;
    return sqrtf(ex * ex + ey * ey + ez * ez);
}

}  // namespace asd

namespace asd {

// Union of the voxels' footprints over the workgroup: wave butterflies, then
// one LDS atomic per wave and bound (rng = a_lo, a_hi, b_lo, b_hi; empty when
// lo > hi).  All threads of the block must call it.
__device__ __forceinline__ void tile_union(bool has, const Footprint &f, int *rng)
{
    int v0 = has ? f.a_lo : (1 << 30), v1 = has ? f.a_hi : -(1 << 30);
    int v2 = has ? f.b_lo : (1 << 30), v3 = has ? f.b_hi : -(1 << 30);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v0 = min(v0, __shfl_xor(v0, off, 64));
        v1 = max(v1, __shfl_xor(v1, off, 64));
        v2 = min(v2, __shfl_xor(v2, off, 64));
        v3 = max(v3, __shfl_xor(v3, off, 64));
    }
    if (threadIdx.x == 0) {
        rng[0] = 1 << 30; rng[1] = -(1 << 30);
        rng[2] = 1 << 30; rng[3] = -(1 << 30);
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&rng[0], v0);
        atomicMax(&rng[1], v1);
        atomicMin(&rng[2], v2);
        atomicMax(&rng[3], v3);
    }
    __syncthreads();
}

constexpr int TILE_MAXA = 256;   // in-plane ray records per chunk
constexpr int TILE_MAXR = 1024;  // (ia, ib) records per chunk

// Tile geometry shared by the kernel and its host check.
struct TileDims {
    int tx, ty, tz;      // tile extents (voxels)
    int ntx, nty, ntz;   // tiles per axis
};

__host__ __device__ __forceinline__ TileDims tile_dims(const ViewGeom &g, int TX, int TY, int TZ)
{
    TileDims d;
    d.tx = TX; d.ty = TY; d.tz = TZ;
    d.ntx = (g.nx + TX - 1) / TX;
    d.nty = (g.ny + TY - 1) / TY;
    d.ntz = (g.nz + TZ - 1) / TZ;
    return d;
}

// Tile-centre coordinates (voxel units, the reference's voxel frame).
__host__ __device__ __forceinline__ void tile_centre(const ViewGeom &g, int x0, int y0, int z0,
                                                     int TX, int TY, int TZ, double &xtc,
                                                     double &ytc, double &ztc)
{
    xtc = (double)x0 + 0.5 * TX - g.nx / 2;
    ytc = (double)y0 + 0.5 * TY - g.ny / 2;
    ztc = ((double)z0 + 0.5 * TZ - g.nz / 2) * g.dz;
}

// One workgroup per TX*TY*TZ tile (TX*TY*TZ == 256).  The rays of the tile's
// footprint are set up once per tile in FP64 and kept in LDS with their
// correction values; each voxel then walks its own footprint (the
// reference's bounds) ia-outer, skipping every ib of an ia whose in-plane
// line misses the voxel.
template <int TX, int TY, int TZ, bool KAHAN, class Epi>
__global__ __launch_bounds__(256) void atx_tile_kernel(ViewGeom g, const float *__restrict__ Y,
                                                       const float *__restrict__ y_det,
                                                       const float *__restrict__ z_det,
                                                       const RayRec *__restrict__ rt, Epi epi)
{
#pragma clang fp contract(off)
    static_assert(TX * TY * TZ == 256, "one voxel per thread");
    __shared__ TileRayXY s_xy[TILE_MAXA];
    __shared__ TileRayZ s_z[TILE_MAXR];
    __shared__ float s_cor[TILE_MAXR];
    __shared__ int s_rng[4];
    const bool stop = epi.skip();
    const TileDims td = tile_dims(g, TX, TY, TZ);
    const unsigned ntiles = (unsigned)td.ntx * td.nty * td.ntz;
    const unsigned t = xcd_block(blockIdx.x, gridDim.x);
    if (t >= ntiles) return;
    const int tile_x = (int)(t % td.ntx), tile_y = (int)((t / td.ntx) % td.nty),
              tile_z = (int)(t / ((unsigned)td.ntx * td.nty));
    const int x0 = tile_x * TX, y0 = tile_y * TY, z0 = tile_z * TZ;
    const int lx = threadIdx.x % TX, ly = (threadIdx.x / TX) % TY, lz = threadIdx.x / (TX * TY);
    const int ix = x0 + lx, iy = y0 + ly, iz = z0 + lz;
    const bool valid = ix < g.nx && iy < g.ny && iz < g.nz;
    const long vox = (long)((size_t)iz * g.ny + iy) * g.nx + ix;
    typename Epi::Pre pre{};
    if (valid) pre = epi.load(vox);
    Footprint f;
    if (valid) f = footprint(g, ix, iy, iz);
    tile_union(valid && f.a_lo <= f.a_hi && f.b_lo <= f.b_hi, f, s_rng);
    const int A0 = s_rng[0], A1 = s_rng[1], B0 = s_rng[2], B1 = s_rng[3];
    double xtc, ytc, ztc;
    tile_centre(g, x0, y0, z0, TX, TY, TZ, xtc, ytc, ztc);
    const float dx = (float)(lx + 0.5 - 0.5 * TX);
    const float dy = (float)(ly + 0.5 - 0.5 * TY);
    const float dzo = (float)((lz + 0.5 - 0.5 * TZ) * g.dz);
    const float hz = g.dz / 2.0f;
    const float sx = g.cos_phi * (-g.SO), sy = g.sin_phi * (-g.SO);
    Acc<KAHAN> acc, accl;
    if (A0 <= A1 && B0 <= B1) {
        const int nB = B1 - B0 + 1;
        int CA = TILE_MAXR / nB;
        if (CA > TILE_MAXA) CA = TILE_MAXA;
        if (CA < 1) CA = 1;
        for (int c0 = A0; c0 <= A1; c0 += CA) {
            const int c1 = c0 + CA - 1 < A1 ? c0 + CA - 1 : A1;
            const int nA = c1 - c0 + 1;
            const int nB_chunk = nB <= TILE_MAXR ? nB : TILE_MAXR;
            // (a single ia with more than TILE_MAXR bins walks ib in pieces below)
            for (int b0 = B0; b0 <= B1; b0 += nB_chunk) {
                const int b1 = b0 + nB_chunk - 1 < B1 ? b0 + nB_chunk - 1 : B1;
                const int nb_ = b1 - b0 + 1;
                __syncthreads();
                for (int k = threadIdx.x; k < nA * nb_; k += 256) {
                    const int ka = k % nA, kb = k / nA;
                    const int ia = c0 + ka, ib = b0 + kb;
                    TileRayXY rxy;
                    TileRayZ rz;
                    tile_ray_setup(g, y_det[ia], z_det[ib], xtc, ytc, ztc, rxy, rz);
                    s_z[k] = rz;
                    s_cor[k] = Y[(long)ib * g.na + ia];
                    if (kb == 0) s_xy[ka] = rxy;
                }
                __syncthreads();
                if (valid) {
                    const int alo = f.a_lo > c0 ? f.a_lo : c0, ahi = f.a_hi < c1 ? f.a_hi : c1;
                    const int blo = f.b_lo > b0 ? f.b_lo : b0, bhi = f.b_hi < b1 ? f.b_hi : b1;
                    for (int ia = alo; ia <= ahi; ia++) {
                        const TileRayXY rxy = s_xy[ia - c0];
                        if (rxy.mode != RAY_X_FIRST && rxy.mode != RAY_Y_FIRST) {
                            // a == 0 or b == 0: the reference's 2-D fallback
                            const float px = sx - f.xc, py = sy - f.yc, pz = 0.0f - f.zc;
                            for (int ib = blo; ib <= bhi; ib++) {
                                const float l = find_l_3d_rec(rt[(long)ib * g.na + ia], px, py,
                                                              pz, hz, g.dz);
                                acc.add(l * s_cor[(ib - b0) * nA + (ia - c0)]);
                                accl.add(l);
                            }
                            continue;
                        }
                        float xy, yx;
                        const unsigned m = tile_inplane(rxy, dx, dy, xy, yx);
                        if (!m) continue;
                        for (int ib = blo; ib <= bhi; ib++) {
                            const int k = (ib - b0) * nA + (ia - c0);
                            const float l = tile_find_l(rxy, s_z[k], m, xy, yx, dx, dy, dzo, hz);
                            acc.add(l * s_cor[k]);
                            accl.add(l);
                        }
                    }
                }
            }
        }
    }
    if (valid && !stop) epi(vox, acc.value(), accl.value(), pre);
}

}  // namespace asd
