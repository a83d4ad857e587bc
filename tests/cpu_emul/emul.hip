// Host-side execution of the projector device functions (test infrastructure).
//
// The geometry helpers in asd_pocs_mex_cuda_demo_amd/csrc/geom_device.hpp are
// __host__ __device__: this file runs the exact functions the GPU kernels
// use, in the kernels' loop order, on the CPU, so tests can compare them with
// the oracle without a GPU.
#include <algorithm>
#include <cstddef>

#include "geom_device.hpp"
#include "projector_kernels.hpp"
#include "projector_tile.hpp"

#include <vector>

using namespace asd;

static ViewGeom mk(float c, float s, float SO, float OD, float scale, float dyd, float dzd,
                   float dz, int nx, int ny, int nz, int na, int nb)
{
    ViewGeom g;
    g.cos_phi = c; g.sin_phi = s; g.SO = SO; g.OD = OD; g.scale = scale; g.dz = dz;
    g.dy_det = dyd; g.dz_det = dzd;
    g.nx = nx; g.ny = ny; g.nz = nz; g.na = na; g.nb = nb;
    geom_finish(g);
    return g;
}

// When set, the flat / dz == 1 shortcuts and the footprint cut are switched
// off (the general expressions over the reference's footprint are
// evaluated), to check that the shortcuts change no bit.
static int g_general = 0;

// When set, the forward projector's pieces are addressed in the wide form
// (element indices) whatever the volume size.
static int g_wide = 0;

static AxRay ray_setup(const ViewGeom &g, float ydet, float zdet)
{
    AxRay r = ax_ray_setup(g, ydet, zdet);
    if (g_general) r.flat = r.dz1 = 0;
    return r;
}

static RayRec ray_rec(const ViewGeom &g, float ydet, float zdet)
{
    RayRec R = make_ray_rec(g, ydet, zdet);
    if (g_general) {
        // the flat shortcut and the two-sided chord off: the product form of
        // the face offsets and the reference's ordered rule
        if (R.mode == RAY_FLAT_X || R.mode == RAY_FLAT_Y)
            R.mode += RAY_X_FIRST_P - RAY_FLAT_X;
        else if (R.mode <= RAY_Y_FIRST)
            R.mode += RAY_X_FIRST_P;
        R.side = 0;
    }
    return R;
}

extern "C" {

void emul_set_general(int on) { g_general = on; }
void emul_set_wide(int on) { g_wide = on; }

// n / d by the kernels' multiply-high division (FastDiv)
void emul_fastdiv(unsigned d, const unsigned *n, unsigned *q, long count)
{
    const FastDiv f = make_fastdiv(d);
    for (long i = 0; i < count; i++) q[i] = fastdiv(n[i], f);
}

// Total (voxel, bin) pairs the backprojector visits for one view: with the
// footprint cut (cut = 1, where the view allows it) or the reference's
// footprint (cut = 0).
long emul_footprint_pairs(float c, float s, const float *ydet, float SO, float OD, float dyd,
                          float dzd, int nx, int ny, int nz, int na, int nb, int cut)
{
    ViewGeom g = mk(c, s, SO, OD, 1.0f, dyd, dzd, 1.0f, nx, ny, nz, na, nb);
    if (cut) g.tight = view_tight(g, ydet);
    long n = 0;
    for (int iz = 0; iz < nz; iz++)
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                Footprint f = footprint(g, ix, iy, iz);
                if (f.a_lo <= f.a_hi && f.b_lo <= f.b_hi)
                    n += (long)(f.a_hi - f.a_lo + 1) * (f.b_hi - f.b_lo + 1);
            }
    return n;
}

// det_coord, the detector grid the forward projector computes in place.
float emul_det_coord(int i, int n, float det) { return det_coord(i, n, det); }

// ax_bundle_slabs against the slabs themselves: over every bundle of K rows
// of the view (rows past nb repeat the last), the number of slabs with a
// piece in the volume outside the bundle's [lo, hi] (must be 0); *kept and
// *total receive the slabs the range keeps and all slabs, summed over
// bundles.
// ax_rows_tile over every block of a row-bundle launch (RPB = 32 bundles of
// 8 lanes per block): returns 1 when it maps [0, nblk) onto itself, and the
// number of blocks it moves in *moved.
int emul_ax_rows_tile_check(int na, int nb, int rows, long *moved)
{
    const int nbb = (nb + rows - 1) / rows;
    const long nbund = (long)na * nbb;
    const unsigned nblk = (unsigned)((nbund * 8 + 255) / 256);
    std::vector<char> hit(nblk, 0);
    *moved = 0;
    for (unsigned b = 0; b < nblk; b++) {
        const unsigned t = ax_rows_tile(b, nblk, na, nb, nbb, 32);
        if (t >= nblk || hit[t]) return 0;
        hit[t] = 1;
        *moved += t != xcd_block(b, nblk);
    }
    return 1;
}

// atx_tile_z_interleave over a launch of per_layer * nlayers tiles in
// xcd_block's order: returns 1 when it maps the tiles onto themselves, and in
// spread[8] the number of distinct eighths of the volume's height each XCD's
// tiles come from.
int emul_atx_tile_z_check(unsigned per_layer, unsigned nlayers, int *spread)
{
    const unsigned n = per_layer * nlayers;
    std::vector<char> hit(n, 0);
    unsigned seen[8] = {0};
    for (unsigned b = 0; b < n; b++) {
        const unsigned t = atx_tile_z_interleave(xcd_block(b, n), per_layer, nlayers);
        if (t >= n || hit[t]) return 0;
        hit[t] = 1;
        if (n >= 16) seen[b % 8] |= 1u << ((t / per_layer) * 8u / nlayers);
    }
    for (int x = 0; x < 8; x++) spread[x] = __builtin_popcount(seen[x]);
    return 1;
}

long emul_ax_slab_range_misses(float c, float s, const float *ydet, const float *zdet, float SO,
                               float OD, float dz, int nx, int ny, int nz, int na, int nb, int K,
                               long *kept, long *total)
{
    ViewGeom g = mk(c, s, SO, OD, 1.0f, 1.0f, 1.0f, dz, nx, ny, nz, na, nb);
    long miss = 0;
    *kept = *total = 0;
    for (int ia = 0; ia < na; ia++)
        for (int ib0 = 0; ib0 < nb; ib0 += K) {
            AxRay r[4];
            float du;
            r[0] = ax_ray_setup(g, ydet[ia], zdet[ib0], du);
            for (int k = 1; k < K; k++) {
                r[k] = r[0];
                ax_ray_axial(r[k], du, zdet[std::min(ib0 + k, nb - 1)]);
            }
            const int nu = r[0].xdriven ? nx : ny;
            int lo, hi;
            ax_bundle_slabs(r, K, nu, lo, hi);
            const SlabAddr A = slab_addr(r[0], nx, 4);
            *total += nu;
            *kept += hi >= lo ? hi - lo + 1 : 0;
            for (int iu = 0; iu < nu; iu++) {
                if (iu >= lo && iu <= hi) continue;
                const SlabT t = ax_slab_t(r[0], A, iu);
                for (int k = 0; k < K; k++) {
                    SlabW w;
                    ax_slab_z(r[k], t, A, ny, w);
                    if (w.ok[0] || w.ok[1] || w.ok[2]) miss++;
                }
            }
        }
    return miss;
}

// The footprint cut against the reference's footprint on nsample voxels of
// one view (drawn with a fixed LCG from seed): returns the number of pairs
// with a nonzero length (find_l_3d_rec, the reference's value) that lie
// outside the cut footprint (must be 0); *kept receives the pairs the cut
// keeps and *nonzero the pairs with a nonzero length.
long emul_footprint_cut_check(float c, float s, const float *ydet, const float *zdet, float SO,
                              float OD, float dyd, float dzd, int nx, int ny, int nz, int na,
                              int nb, int nsample, unsigned seed, long *kept, long *nonzero)
{
    ViewGeom g = mk(c, s, SO, OD, 1.0f, dyd, dzd, 1.0f, nx, ny, nz, na, nb);
    const int tight = view_tight(g, ydet);
    long miss = 0;
    *kept = *nonzero = 0;
    for (int k = 0; k < nsample; k++) {
        seed = seed * 1664525u + 1013904223u;
        const int ix = (int)((seed >> 8) % (unsigned)nx);
        seed = seed * 1664525u + 1013904223u;
        const int iy = (int)((seed >> 8) % (unsigned)ny);
        seed = seed * 1664525u + 1013904223u;
        const int iz = (int)((seed >> 8) % (unsigned)nz);
        g.tight = 0;
        const Footprint fr = footprint(g, ix, iy, iz);
        g.tight = tight;
        const Footprint fc = footprint(g, ix, iy, iz);
        if (fc.a_lo <= fc.a_hi && fc.b_lo <= fc.b_hi)
            *kept += (long)(fc.a_hi - fc.a_lo + 1) * (fc.b_hi - fc.b_lo + 1);
        const float px = g.cos_phi * (-g.SO) - fr.xc, py = g.sin_phi * (-g.SO) - fr.yc,
                    pz = 0.0f - fr.zc;
        for (int ib = fr.b_lo; ib <= fr.b_hi; ib++)
            for (int ia = fr.a_lo; ia <= fr.a_hi; ia++) {
                const RayRec R = make_ray_rec(g, ydet[ia], zdet[ib]);
                if (find_l_3d_rec(R, px, py, pz, g.dz / 2.0f, g.dz) == 0.0f) continue;
                ++*nonzero;
                if (ia < fc.a_lo || ia > fc.a_hi || ib < fc.b_lo || ib > fc.b_hi) miss++;
            }
    }
    return miss;
}

// view_general for a view (after view_tight, as the plans set them).
int emul_view_general(float c, float s, const float *ydet, const float *zdet, float SO, float OD,
                      float dyd, float dzd, float dz, int nx, int ny, int nz, int na, int nb)
{
    ViewGeom g = mk(c, s, SO, OD, 1.0f, dyd, dzd, dz, nx, ny, nz, na, nb);
    g.tight = view_tight(g, ydet);
    return view_general(g, ydet, zdet);
}

// Lengths of n (ray, voxel) pairs by the backprojector's record path
// (make_ray_rec, find_l_3d_rec): pair k is the ray of the view (c[k], s[k],
// SO[k], OD[k]) to detector point (ydet[k], zdet[k]) and the unit voxel
// centred at (xc[k], yc[k], zc[k]) in a volume of nz slices.
void emul_find_l_3d_pairs(long n, const float *c, const float *s, const float *SO,
                          const float *OD, const float *ydet, const float *zdet, const float *xc,
                          const float *yc, const float *zc, int nz, float *out)
{
    for (long k = 0; k < n; k++) {
        ViewGeom g = mk(c[k], s[k], SO[k], OD[k], 1.0f, 1.0f, 1.0f, 1.0f, 4, 4, nz, 4, 4);
        const RayRec R = ray_rec(g, ydet[k], zdet[k]);
        const float px = g.cos_phi * (-g.SO) - xc[k], py = g.sin_phi * (-g.SO) - yc[k],
                    pz = 0.0f - zc[k];
        out[k] = find_l_3d_rec(R, px, py, pz, g.dz / 2.0f, g.dz);
    }
}

}  // extern "C"

// Lane partials of one ray combined in the order of group_sum (quad swaps,
// half-row and row mirrors, then the 16- and 32-lane swaps, within groups of
// at most 64 lanes), then the wave partials in wave order; compensated
// partials in FP64 as on the device.
static void combine_lanes(double *acc, double *accl, int lpr, int kahan, double &a, double &l)
{
    const int wl = lpr < 64 ? lpr : 64;
    for (int w = 2; w <= wl; w *= 2) {
        double na_[256], nl_[256];
        for (int lane = 0; lane < lpr; lane++) {
            const int o = w == 8    ? (lane & ~7) | (7 - (lane & 7))
                          : w == 16 ? (lane & ~15) | (15 - (lane & 15))
                                    : lane ^ (w / 2);
            na_[lane] = kahan ? acc[lane] + acc[o] : (double)((float)acc[lane] + (float)acc[o]);
            nl_[lane] = kahan ? accl[lane] + accl[o] : (double)((float)accl[lane] + (float)accl[o]);
        }
        for (int lane = 0; lane < lpr; lane++) {
            acc[lane] = na_[lane];
            accl[lane] = nl_[lane];
        }
    }
    a = acc[0];
    l = accl[0];
    for (int w = 1; w * 64 < lpr; w++) {
        a = kahan ? a + acc[w * 64] : (double)((float)a + (float)acc[w * 64]);
        l = kahan ? l + accl[w * 64] : (double)((float)l + (float)accl[w * 64]);
    }
}

template <bool KAHAN>
static void lane_partial(const Acc<KAHAN> &a, const Acc<KAHAN> &l, double &pa, double &pl)
{
    pa = KAHAN ? (double)a.s - (double)a.c : (double)a.s;
    pl = KAHAN ? (double)l.s - (double)l.c : (double)l.s;
}

// rows == 1: ax_kernel<LPR=lpr>, lane j of a ray takes slabs j, j+lpr, ...
// rows > 1: ax_rows_kernel<LPR=lpr, K=rows>, the same lanes over bundles of
// rows detector rows of one column, each slab's transverse half shared.
template <bool KAHAN>
static void emul_ax_t(const float *X, float *y, float *sr, const ViewGeom &g, const float *ydet,
                      const float *zdet, float scale, int lpr, int rows)
{
    const int nx = g.nx, ny = g.ny, na = g.na, nb = g.nb;
    static double acc[8][256], accl[8][256];
    for (int ib0 = 0; ib0 < nb; ib0 += rows)
        for (int ia = 0; ia < na; ia++) {
            float du;
            AxRay r[8];
            r[0] = ax_ray_setup(g, ydet[ia], zdet[ib0], du);
            for (int k = 1; k < rows; k++) {
                r[k] = r[0];
                ax_ray_axial(r[k], du, zdet[std::min(ib0 + k, nb - 1)]);
            }
            if (g_general)
                for (int k = 0; k < rows; k++) r[k].flat = r[k].dz1 = 0;
            // the kernels' march for views without flat rays (under g_general
            // every ray takes it, flat ones included)
            const bool noflat = view_has_no_flat_ray(g) || g_general;
            const int nu = r[0].xdriven ? nx : ny;
            // the buffer form's byte offsets where the kernels use it, else
            // element indices (the same voxels)
            const bool buf = !g_wide && vol_buffer_ok((size_t)nx * ny * g.nz, nx);
            const VolRead<false> Vb(X, (size_t)nx * ny * g.nz);
            const VolRead<true> Vw(X, (size_t)nx * ny * g.nz);
            const SlabAddr A = slab_addr(r[0], nx, buf ? VolRead<false>::unit : VolRead<true>::unit);
            for (int lane = 0; lane < lpr; lane++) {
                Acc<KAHAN> a[8], l[8];
                if (rows == 1) {
                    // flat rays take the kernels' flat-specialised slab code
                    for (int iu0 = lane; iu0 < nu; iu0 += 4 * lpr)
                        if (r[0].flat && buf)
                            ax_slabs<4, KAHAN, false, VolRead<false>, true>(r[0], A, Vb, iu0, lpr,
                                                                            nu, ny, a[0], l[0]);
                        else if (r[0].flat)
                            ax_slabs<4, KAHAN, false, VolRead<true>, true>(r[0], A, Vw, iu0, lpr,
                                                                           nu, ny, a[0], l[0]);
                        else if (noflat && buf)
                            ax_slabs<4, KAHAN, false, VolRead<false>, FLAT_NONE>(
                                r[0], A, Vb, iu0, lpr, nu, ny, a[0], l[0]);
                        else if (noflat)
                            ax_slabs<4, KAHAN, false, VolRead<true>, FLAT_NONE>(
                                r[0], A, Vw, iu0, lpr, nu, ny, a[0], l[0]);
                        else if (buf)
                            ax_slabs<4, KAHAN, false>(r[0], A, Vb, iu0, lpr, nu, ny, a[0], l[0]);
                        else
                            ax_slabs<4, KAHAN, false>(r[0], A, Vw, iu0, lpr, nu, ny, a[0], l[0]);
                } else {
                    for (int iu = lane; iu < nu; iu += lpr) {
                        const SlabT t = ax_slab_t(r[0], A, iu);
                        for (int k = 0; k < rows; k++) {
                            SlabW s;
                            if (noflat)
                                ax_slab_z<FLAT_NONE>(r[k], t, A, ny, s);
                            else
                                ax_slab_z(r[k], t, A, ny, s);
                            for (int p = 0; p < 3; p++)
                                if (s.ok[p]) {
                                    const float xv = buf ? Vb(s.off[p], true) : Vw(s.off[p], true);
                                    a[k].add(s.wl[p] * xv);
                                    l[k].add(s.wl[p]);
                                }
                        }
                    }
                }
                for (int k = 0; k < rows; k++) lane_partial(a[k], l[k], acc[k][lane], accl[k][lane]);
            }
            for (int k = 0; k < rows && ib0 + k < nb; k++) {
                double sa, sl;
                combine_lanes(acc[k], accl[k], lpr, KAHAN, sa, sl);
                const long id = (long)(ib0 + k) * na + ia;
                y[id] = (float)sa * scale;
                sr[id] = (float)sl * scale;
            }
        }
}

extern "C" {

void emul_ax(const float *X, float *y, float *sr, float c, float s, const float *ydet,
             const float *zdet, float SO, float OD, float scale, float dz, int nx, int ny, int nz,
             int na, int nb, int lpr, int kahan, int rows)
{
    ViewGeom g = mk(c, s, SO, OD, scale, 1.0f, 1.0f, dz, nx, ny, nz, na, nb);
    if (rows < 1 || rows > 8 || lpr < 1 || lpr > 256) return;
    if (kahan)
        emul_ax_t<true>(X, y, sr, g, ydet, zdet, scale, lpr, rows);
    else
        emul_ax_t<false>(X, y, sr, g, ydet, zdet, scale, lpr, rows);
}

// atx_kernel: one voxel at a time, footprint ib outer / ia inner.
void emul_atx(float *X, const float *y, float *sc, float c, float s, const float *ydet,
              const float *zdet, float SO, float OD, float scale, float dyd, float dzd, float dz,
              int nx, int ny, int nz, int na, int nb)
{
    ViewGeom g = mk(c, s, SO, OD, scale, dyd, dzd, dz, nx, ny, nz, na, nb);
    if (!g_general) g.tight = view_tight(g, ydet);
    if (!g_general) g.general = view_general(g, ydet, zdet);
    long M = (long)na * nb;
    RayRec *rt = new RayRec[M];
    for (long id = 0; id < M; id++) rt[id] = ray_rec(g, ydet[id % na], zdet[id / na]);
    for (long v = 0; v < (long)nx * ny * nz; v++) {
        int ix = (int)(v % nx);
        long r = v / nx;
        int iy = (int)(r % ny), iz = (int)(r / ny);
        Footprint f = footprint(g, ix, iy, iz);
        float sx = g.cos_phi * (-g.SO), sy = g.sin_phi * (-g.SO);
        float px = sx - f.xc, py = sy - f.yc, pz = 0.0f - f.zc;
        float hz = g.dz / 2.0f;
        Acc<false> acc, accl;
        if (g.general) {
            // the kernels' walk for views of general rays (two pairs per load batch)
            if (f.a_lo <= f.a_hi && f.b_lo <= f.b_hi)
                atx_voxel_walk_general<false, 2>(g, BinsPtr{rt, y}, f, px, py, pz, acc, accl);
            X[v] = acc.value() * scale;
            sc[v] = accl.value() * scale;
            continue;
        }
        for (int ib = f.b_lo; ib <= f.b_hi; ib++)
            for (int ia = f.a_lo; ia <= f.a_hi; ia++) {
                const RayRec &R = rt[(long)ib * na + ia];
                if (!g.tight && !inplane_may_hit(R, px, py)) continue;
                float l = find_l_3d_rec(R, px, py, pz, hz, g.dz);
                acc.add(l * y[(long)ib * na + ia]);
                accl.add(l);
            }
        X[v] = acc.value() * scale;
        sc[v] = accl.value() * scale;
    }
    delete[] rt;
}

// atx_tile_kernel (precision variant) tile by tile: the tile's union
// footprint, FP64 tile-local ray records, then every voxel ia-outer with the
// in-plane skip.  tx, ty, tz as the kernel's template arguments.
void emul_atx_prec(float *X, const float *y, float *sc, float c, float s, const float *ydet,
                   const float *zdet, float SO, float OD, float scale, float dyd, float dzd,
                   float dz, int nx, int ny, int nz, int na, int nb, int TX, int TY, int TZ)
{
    ViewGeom g = mk(c, s, SO, OD, scale, dyd, dzd, dz, nx, ny, nz, na, nb);
    if (!g_general) g.tight = view_tight(g, ydet);
    long M = (long)na * nb;
    std::vector<RayRec> rt(M);
    for (long id = 0; id < M; id++) rt[id] = ray_rec(g, ydet[id % na], zdet[id / na]);
    TileDims td = tile_dims(g, TX, TY, TZ);
    for (int tzi = 0; tzi < td.ntz; tzi++)
        for (int tyi = 0; tyi < td.nty; tyi++)
            for (int txi = 0; txi < td.ntx; txi++) {
                const int x0 = txi * TX, y0 = tyi * TY, z0 = tzi * TZ;
                int A0 = 1 << 30, A1 = -(1 << 30), B0 = 1 << 30, B1 = -(1 << 30);
                for (int lz = 0; lz < TZ; lz++)
                    for (int ly = 0; ly < TY; ly++)
                        for (int lx = 0; lx < TX; lx++) {
                            int ix = x0 + lx, iy = y0 + ly, iz = z0 + lz;
                            if (ix >= nx || iy >= ny || iz >= nz) continue;
                            Footprint f = footprint(g, ix, iy, iz);
                            if (f.a_lo > f.a_hi || f.b_lo > f.b_hi) continue;
                            A0 = f.a_lo < A0 ? f.a_lo : A0; A1 = f.a_hi > A1 ? f.a_hi : A1;
                            B0 = f.b_lo < B0 ? f.b_lo : B0; B1 = f.b_hi > B1 ? f.b_hi : B1;
                        }
                double xtc, ytc, ztc;
                tile_centre(g, x0, y0, z0, TX, TY, TZ, xtc, ytc, ztc);
                int nA = A1 - A0 + 1, nB = B1 - B0 + 1;
                std::vector<TileRayXY> rxy(nA > 0 ? nA : 0);
                std::vector<TileRayZ> rz((nA > 0 && nB > 0) ? (size_t)nA * nB : 0);
                for (int kb = 0; kb < nB; kb++)
                    for (int ka = 0; ka < nA; ka++) {
                        TileRayXY a;
                        tile_ray_setup(g, ydet[A0 + ka], zdet[B0 + kb], xtc, ytc, ztc, a,
                                       rz[(size_t)kb * nA + ka]);
                        if (kb == 0) rxy[ka] = a;
                    }
                for (int lz = 0; lz < TZ; lz++)
                    for (int ly = 0; ly < TY; ly++)
                        for (int lx = 0; lx < TX; lx++) {
                            int ix = x0 + lx, iy = y0 + ly, iz = z0 + lz;
                            if (ix >= nx || iy >= ny || iz >= nz) continue;
                            Footprint f = footprint(g, ix, iy, iz);
                            float dxo = (float)(lx + 0.5 - 0.5 * TX), dyo = (float)(ly + 0.5 - 0.5 * TY);
                            float dzo = (float)((lz + 0.5 - 0.5 * TZ) * g.dz), hz = g.dz / 2.0f;
                            float sx = g.cos_phi * (-g.SO), sy = g.sin_phi * (-g.SO);
                            Acc<true> acc, accl;
                            for (int ia = f.a_lo; ia <= f.a_hi; ia++) {
                                const TileRayXY &r = rxy[ia - A0];
                                if (r.mode != RAY_X_FIRST && r.mode != RAY_Y_FIRST) {
                                    float px = sx - f.xc, py = sy - f.yc, pz = 0.0f - f.zc;
                                    for (int ib = f.b_lo; ib <= f.b_hi; ib++) {
                                        float l = find_l_3d_rec(rt[(long)ib * na + ia], px, py, pz, hz, g.dz);
                                        acc.add(l * y[(long)ib * na + ia]);
                                        accl.add(l);
                                    }
                                    continue;
                                }
                                float xy, yx;
                                unsigned m = tile_inplane(r, dxo, dyo, xy, yx);
                                if (!m) continue;
                                for (int ib = f.b_lo; ib <= f.b_hi; ib++) {
                                    float l = tile_find_l(r, rz[(size_t)(ib - B0) * nA + (ia - A0)], m, xy,
                                                          yx, dxo, dyo, dzo, hz);
                                    acc.add(l * y[(long)ib * na + ia]);
                                    accl.add(l);
                                }
                            }
                            long v = ((long)iz * ny + iy) * nx + ix;
                            X[v] = acc.value() * scale;
                            sc[v] = accl.value() * scale;
                        }
            }
}

}  // extern "C"
