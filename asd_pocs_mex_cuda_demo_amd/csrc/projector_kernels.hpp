// Projector kernel templates shared by the plain operators and the fused
// SART stages.  The epilogue functor decides what a finished ray sum or voxel
// sum becomes (plain y/sr, or the SART correction / image update).
#pragma once

#include <type_traits>

#include "geom_device.hpp"
#include "projector_tile.hpp"

namespace asd {

// The stop flag of the device-driven loop is read at kernel entry but only
// consulted before the epilogue's stores, so its load overlaps the work
// instead of delaying it (a stopped loop's kernels compute and store nothing;
// the host stops replaying within a few iterations).
//
// Forward projector: LPR lanes march one ray together (lane j takes slabs
// j, j+LPR, ...), then a butterfly over the lanes of a wave and, when a ray
// spans several waves (LPR = 128, 256), a fixed-order sum of the wave
// partials through LDS.  With LPR = 1 a thread walks its ray in slab order
// like the reference kernel (cu:375, :576); a 2-D view (a few hundred rays)
// gets up to four waves per ray so it still fills the chip, and lanes
// walking along x read consecutive voxels.
template <int LPR, bool KAHAN, bool WIDE, class Epi>
__global__ __launch_bounds__(256) void ax_kernel(ViewGeom g, const float *__restrict__ X,
                                                 const float *__restrict__ y_det,
                                                 const float *__restrict__ z_det, Epi epi)
{
    constexpr int RPB = LPR >= 256 ? 1 : 256 / LPR;  // rays per block
    constexpr int WL = LPR < 64 ? LPR : 64;          // lanes of a ray inside one wave
    constexpr int WPR = LPR > 64 ? LPR / 64 : 1;     // waves per ray
    const bool stop = epi.skip();
    const long nrays = (long)g.na * g.nb;
    const int lane = threadIdx.x % LPR;
    // XCD-contiguous rays: neighbouring rays of a fan cross the same voxels,
    // so each XCD's L2 holds one wedge of the volume instead of all of it
    const long id = (long)xcd_block(blockIdx.x, gridDim.x) * RPB + threadIdx.x / LPR;
    const bool active = id < nrays;
    const VolRead<WIDE> V(X, (size_t)g.nx * g.ny * g.nzs);  // the window's slices
    typename Epi::Pre pre{};
    if (active && lane == 0) pre = epi.load(id);
    Acc<KAHAN> sa, sl;
    if (active) {
        // id < na * nb < 2^31 and ib < nb < 2^24 (index_fits)
        const unsigned ub = fastdiv((unsigned)id, g.div_na);
        const int ia = (int)((unsigned)id - mad24(ub, (unsigned)g.na, 0u)), ib = (int)ub;
        // null tables: the gateway's regular detector grid, computed in place
        const float yd = y_det ? y_det[ia] : det_coord(ia, g.na, g.dy_det);
        const float zd = z_det ? z_det[ib] : det_coord(ib, g.nb, g.dz_det);
        AxRay r = ax_ray_setup(g, yd, zd);
        const SlabAddr A = slab_addr(r, g.nx, VolRead<WIDE>::unit);
        int nu = r.xdriven ? g.nx : g.ny;
        // slabs per load batch: two when a ray has 256 lanes, else four (a
        // lane then walks its slabs in one or two batches)
        constexpr int U = LPR >= 256 ? 2 : 4;
        auto march = [&](auto ft, auto flat) {
            for (int iu0 = lane; iu0 < nu; iu0 += U * LPR)
                ax_slabs<U, KAHAN, decltype(ft)::value, VolRead<WIDE>, decltype(flat)::value>(
                    r, A, V, iu0, LPR, nu, g.ny, sa, sl);
        };
        // fast_t depends on the view only, every ray of a 2-D view is flat,
        // and no ray of a view of several slices is: one branch for the
        // whole march
        if (r.fast_t && r.flat)
            march(std::true_type{}, std::integral_constant<int, FLAT_ALL>{});
        else if (r.fast_t && view_has_no_flat_ray(g))
            march(std::true_type{}, std::integral_constant<int, FLAT_NONE>{});
        else if (r.fast_t)
            march(std::true_type{}, std::integral_constant<int, FLAT_ANY>{});
        else
            march(std::false_type{}, std::integral_constant<int, FLAT_ANY>{});
    }
    // compensated partials are combined in FP64, plain ones in FP32
    using R = typename std::conditional<KAHAN, double, float>::type;
    R a = KAHAN ? (R)sa.s - (R)sa.c : (R)sa.s;
    R b = KAHAN ? (R)sl.s - (R)sl.c : (R)sl.s;
    if constexpr (WL > 1) {
        a = group_sum<WL>(a);
        b = group_sum<WL>(b);
    }
    if constexpr (WPR > 1) {
        // RPB rays per block, WPR waves per ray: wave w of the block holds
        // the partial of ray w / WPR; lane 0 of each ray adds them in order
        __shared__ R sp[RPB * WPR][2];
        const int w = threadIdx.x / 64;
        if ((threadIdx.x & 63) == 0) {
            sp[w][0] = a;
            sp[w][1] = b;
        }
        __syncthreads();
        if (lane == 0) {
            const int w0 = (threadIdx.x / LPR) * WPR;
            a = sp[w0][0];
            b = sp[w0][1];
#pragma unroll
            for (int k = 1; k < WPR; k++) {
                a += sp[w0 + k][0];
                b += sp[w0 + k][1];
            }
        }
    }
    if (active && lane == 0 && !stop) epi(id, (float)a, (float)b, pre);
}

// The tile (block of bundles) a workgroup of the row-bundle forward projector
// takes.  xcd_block gives each XCD a contiguous band of row bundles.  On views
// of at least 2^19 rays, whose bands divide evenly into column blocks of RPB
// bundles:
//  - each XCD's row bundles are 8 blocks spread over the detector's height
//    (when 64 blocks divide them evenly), since the rays of the top and
//    bottom rows leave the volume early and a band of them is less work;
//  - an XCD's tiles are taken column block by column block from the centre
//    of the detector outwards, all of its row bundles for each, so the
//    longest rays (those through the middle of the volume) start first and
//    the short ones at the detector's edges fill the end of the launch.
// The map is a bijection on [0, ncb * nbb) (checked on the CPU); otherwise
// the tile is xcd_block's.  512^3 view with 1024 x 768 bins: 727 -> 684 us
// (longest first), 683 -> 661 us (spread over the height); 256^3 with
// 384 x 384 bins ran 1% slower with the first, and keeps the plain order.
__host__ __device__ __forceinline__ unsigned ax_rows_tile(unsigned b, unsigned nblk, int na, int nb,
                                                          int nbb, int rpb)
{
    const unsigned tile = xcd_block(b, nblk);
    const unsigned ncb = (unsigned)na / (unsigned)rpb, ubb = (unsigned)nbb;
    if ((long)na * nb < (1L << 19) || (unsigned)na % (unsigned)rpb != 0 || ubb % 8u != 0 ||
        nblk != ncb * ubb || ncb < 2)
        return tile;
    const unsigned nrbx = ubb / 8u, x = tile / (ncb * nrbx);
    const unsigned k = tile - x * ncb * nrbx;
    const unsigned ci = k / nrbx, r = k - ci * nrbx;
    const unsigned half = ncb / 2u;
    const unsigned cb = (ci & 1u) ? half - 1u - ci / 2u : half + ci / 2u;
    // the XCD's row bundles: blocks x, x + 8, ... of 64 blocks over the
    // detector's height when they divide evenly (the rays of the top and
    // bottom rows leave the volume early), else its own band
    unsigned rb = x * nrbx + r;
    if (ubb % 64u == 0) {
        const unsigned kb = ubb / 64u, i = r / kb, lk = r - i * kb;
        rb = (i * 8u + x) * kb + lk;
    }
    return rb * ncb + cb;
}

// Forward projector over bundles of K rays that share a detector column (ia)
// and take consecutive rows ib0 .. ib0 + K - 1.  Their in-plane lines are the
// same, so each slab's transverse half (ax_slab_t: extent, cells, crossing
// fraction) is computed once for the K rays and only the axial half
// (ax_slab_z) per ray; every ray's pieces, lengths and sums are the ones
// ax_kernel forms for it.  LPR lanes march the bundle together as above; the
// K ray sums are reduced the same way and lane k stores ray k.  Rows past nb
// repeat the last row and are not stored.  A lane takes U slabs per load
// batch (iu, iu + LPR, ...), so 3 * U * K loads are in flight together.  The
// march covers only the slabs ax_bundle_slabs leaves (a 512^3 volume under a
// 1024 x 768 detector puts about a quarter of the slabs outside it).
template <int LPR, int K, int U, bool KAHAN, bool WIDE, class Epi>
__global__ __launch_bounds__(256) void ax_rows_kernel(ViewGeom g, const float *__restrict__ X,
                                                      const float *__restrict__ y_det,
                                                      const float *__restrict__ z_det, Epi epi)
{
#pragma clang fp contract(off)
    static_assert(LPR >= K, "lane k stores row k");
    constexpr int RPB = LPR >= 256 ? 1 : 256 / LPR;  // bundles per block
    constexpr int WL = LPR < 64 ? LPR : 64;
    constexpr int WPR = LPR > 64 ? LPR / 64 : 1;
    const bool stop = epi.skip();
    const int nbb = (g.nb + K - 1) / K;  // bundles per detector column
    const long nbund = (long)g.na * nbb;
    const int lane = threadIdx.x % LPR;
    const long id = (long)ax_rows_tile(blockIdx.x, gridDim.x, g.na, g.nb, nbb, RPB) * RPB +
                    threadIdx.x / LPR;
    const bool active = id < nbund;
    // id < na * nb < 2^31 (index_fits)
    const unsigned ub = active ? fastdiv((unsigned)id, g.div_na) : 0u;
    const int ia = (int)((unsigned)id - mad24(ub, (unsigned)g.na, 0u)), ib0 = (int)ub * K;
    const bool mine = active && lane < K && ib0 + lane < g.nb;
    const VolRead<WIDE> V(X, (size_t)g.nx * g.ny * g.nzs);  // the window's slices
    const long ray = (long)(ib0 + lane) * g.na + ia;
    typename Epi::Pre pre{};
    if (mine) pre = epi.load(ray);
    Acc<KAHAN> sa[K], sl[K];
    if (active) {
        const float yd = y_det ? y_det[ia] : det_coord(ia, g.na, g.dy_det);
        float du;
        AxRay r[K];
        r[0] = ax_ray_setup(g, yd, z_det ? z_det[ib0] : det_coord(ib0, g.nb, g.dz_det), du);
#pragma unroll
        for (int k = 1; k < K; k++) {
            const int ib = min(ib0 + k, g.nb - 1);
            r[k] = r[0];
            ax_ray_axial(r[k], du, z_det ? z_det[ib] : det_coord(ib, g.nb, g.dz_det));
        }
        const int nu = r[0].xdriven ? g.nx : g.ny;
        // the slabs that may hold a piece in the volume; lane j still takes
        // the slabs iu == j (mod LPR), so its partial sums keep their terms
        int lo, hi;
        ax_bundle_slabs(r, K, nu, lo, hi);
        const int first = lo + ((lane - lo) % LPR + LPR) % LPR;
        const SlabAddr A = slab_addr(r[0], g.nx, VolRead<WIDE>::unit);
        auto march = [&](auto ft, auto flat) {
            for (int iu0 = first; iu0 <= hi; iu0 += U * LPR) {
                SlabW s[U][K];
                float xv[U][K][3];
#pragma unroll
                for (int j = 0; j < U; j++) {
                    const int iu = iu0 + j * LPR;
                    if (iu <= hi) {
                        const SlabT t = ax_slab_t<decltype(ft)::value>(r[0], A, iu);
#pragma unroll
                        for (int k = 0; k < K; k++)
                            ax_slab_z<decltype(flat)::value>(r[k], t, A, g.ny, s[j][k]);
                    } else {
#pragma unroll
                        for (int k = 0; k < K; k++)
#pragma unroll
                            for (int p = 0; p < 3; p++) {
                                s[j][k].ok[p] = false;
                                s[j][k].off[p] = VOL_NONE;
                                s[j][k].wl[p] = unspecified(0.0f);
                            }
                    }
                }
#pragma unroll
                for (int j = 0; j < U; j++)
#pragma unroll
                    for (int k = 0; k < K; k++)
#pragma unroll
                        for (int p = 0; p < 3; p++) xv[j][k][p] = V(s[j][k].off[p], s[j][k].ok[p]);
                // each row's pieces in slab order
#pragma unroll
                for (int j = 0; j < U; j++)
#pragma unroll
                    for (int k = 0; k < K; k++)
#pragma unroll
                        for (int p = 0; p < 3; p++)
                            if (s[j][k].ok[p]) {
                                sa[k].add(s[j][k].wl[p] * xv[j][k][p]);
                                sl[k].add(s[j][k].wl[p]);
                            }
            }
        };
        // fast_t depends on the view only, and no ray of a view of several
        // slices is flat: one branch for the whole march
        if (r[0].fast_t && view_has_no_flat_ray(g))
            march(std::true_type{}, std::integral_constant<int, FLAT_NONE>{});
        else if (r[0].fast_t)
            march(std::true_type{}, std::integral_constant<int, FLAT_ANY>{});
        else
            march(std::false_type{}, std::integral_constant<int, FLAT_ANY>{});
    }
    using R = typename std::conditional<KAHAN, double, float>::type;
    R a[K], b[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        a[k] = KAHAN ? (R)sa[k].s - (R)sa[k].c : (R)sa[k].s;
        b[k] = KAHAN ? (R)sl[k].s - (R)sl[k].c : (R)sl[k].s;
        if constexpr (WL > 1) {
            a[k] = group_sum<WL>(a[k]);
            b[k] = group_sum<WL>(b[k]);
        }
    }
    // the butterfly leaves the group's sum in every lane; lane k of each
    // wave keeps row k
    const int wl = threadIdx.x % WL;
    R am = a[0], bm = b[0];
#pragma unroll
    for (int k = 1; k < K; k++) {
        am = wl == k ? a[k] : am;
        bm = wl == k ? b[k] : bm;
    }
    if constexpr (WPR > 1) {
        __shared__ R sp[RPB * WPR][2 * K];
        const int w = threadIdx.x / 64;
        if (wl < K) {
            sp[w][2 * wl] = am;
            sp[w][2 * wl + 1] = bm;
        }
        __syncthreads();
        if (lane < K) {
            const int w0 = (threadIdx.x / LPR) * WPR;
            am = sp[w0][2 * lane];
            bm = sp[w0][2 * lane + 1];
#pragma unroll
            for (int j = 1; j < WPR; j++) {
                am += sp[w0 + j][2 * lane];
                bm += sp[w0 + j][2 * lane + 1];
            }
        }
    }
    if (mine && !stop) epi(ray, (float)am, (float)bm, pre);
}

// Forward projector that sums every ray in slab order, the order of the
// reference kernel, where one thread per ray adds slab after slab into
// y[id] and sr[id] (cu:375-772): one wave per ray.  Lane j forms the pieces
// of slab c + j of each run of 64 slabs (the products w*l*x and weights w*l
// of ax_slabs); then the wave adds the run's terms in slab order and, per
// slab, in piece order, reading each lane's terms as wave-uniform values.  A
// piece outside the volume adds +0, which leaves every sum unchanged (a sum
// that starts at +0 never becomes -0).  The sums are therefore the
// reference's bit for bit, while a run's geometry and loads are spread over
// the wave.  Used where results must follow the reference's order exactly
// (the reference-arithmetic plan and the slab-ordered drop-in form).
template <bool WIDE, class Epi>
__global__ __launch_bounds__(256) void ax_ordered_kernel(ViewGeom g, const float *__restrict__ X,
                                                         const float *__restrict__ y_det,
                                                         const float *__restrict__ z_det, Epi epi)
{
#pragma clang fp contract(off)
    const bool stop = epi.skip();
    const long nrays = (long)g.na * g.nb;
    const int lane = threadIdx.x & 63;
    const long id = (long)xcd_block(blockIdx.x, gridDim.x) * 4 + threadIdx.x / 64;
    if (id >= nrays) return;  // whole waves
    const VolRead<WIDE> V(X, (size_t)g.nx * g.ny * g.nzs);  // the window's slices
    typename Epi::Pre pre{};
    if (lane == 0) pre = epi.load(id);
    // id < na * nb < 2^31 and ib < nb < 2^24 (index_fits)
    const unsigned ub = fastdiv((unsigned)id, g.div_na);
    const int ia = (int)((unsigned)id - mad24(ub, (unsigned)g.na, 0u)), ib = (int)ub;
    const float yd = y_det ? y_det[ia] : det_coord(ia, g.na, g.dy_det);
    const float zd = z_det ? z_det[ib] : det_coord(ib, g.nb, g.dz_det);
    const AxRay r = ax_ray_setup(g, yd, zd);
    const SlabAddr A = slab_addr(r, g.nx, VolRead<WIDE>::unit);
    const int nu = r.xdriven ? g.nx : g.ny;
    float acc = 0.0f, accl = 0.0f;  // wave-uniform running sums
    auto march = [&](auto ft, auto flat) {
        for (int c0 = 0; c0 < nu; c0 += 64) {
            const int iu = c0 + lane;
            SlabW sw;
            if (iu < nu) {
                ax_slab_geom<decltype(ft)::value, decltype(flat)::value>(r, A, iu, g.ny, sw);
            } else {
                for (int p = 0; p < 3; p++) {
                    sw.ok[p] = false;
                    sw.off[p] = VOL_NONE;
                    sw.wl[p] = unspecified(0.0f);
                }
            }
            float t[3], tl[3];
#pragma unroll
            for (int p = 0; p < 3; p++) {
                const float xv = V(sw.off[p], sw.ok[p]);
                t[p] = sw.ok[p] ? sw.wl[p] * xv : 0.0f;
                tl[p] = sw.ok[p] ? sw.wl[p] : 0.0f;
            }
            // only slabs with a piece in the volume add anything
            unsigned long long live = __ballot(sw.ok[0] || sw.ok[1] || sw.ok[2]);
            while (live) {
                const int l = __builtin_ctzll(live);
                live &= live - 1;
#pragma unroll
                for (int p = 0; p < 3; p++) {
                    acc += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t[p]), l));
                    accl += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tl[p]), l));
                }
            }
        }
    };
    if (r.fast_t && r.flat)
        march(std::true_type{}, std::true_type{});
    else if (r.fast_t)
        march(std::true_type{}, std::false_type{});
    else
        march(std::false_type{}, std::false_type{});
    if (lane == 0 && !stop) epi(id, acc, accl, pre);
}

// Slab-order forward projector launch for one view (4 rays per block).
template <class Epi>
int dispatch_ax_ordered(const ViewGeom &g, const float *x, const float *y_det,
                        const float *z_det, const Epi &epi, hipStream_t s)
{
    const long nrays = (long)g.na * g.nb;
    const unsigned nblk = (unsigned)((nrays + 3) / 4);
    if (vol_buffer_ok((size_t)g.nx * g.ny * g.nzs, g.nx))
        klaunch(ax_ordered_kernel<false, Epi>, nblk, 256, 0, s, g, x, y_det, z_det, epi);
    else
        klaunch(ax_ordered_kernel<true, Epi>, nblk, 256, 0, s, g, x, y_det, z_det, epi);
    ASD_LAUNCH_CHECK("ax_ordered_kernel");
    return ASD_OK;
}

template <bool KAHAN, bool WIDE, class Epi>
int dispatch_ax_t(AxShape sh, const ViewGeom &g, const float *x, const float *y_det,
                const float *z_det, const Epi &epi, hipStream_t s)
{
    if (sh.rows > 1) {
        const long nbund = (long)g.na * ((g.nb + sh.rows - 1) / sh.rows);
        const unsigned nblk = (unsigned)((nbund * sh.lpr + 255) / 256);
        if (sh.lpr == 8 && sh.rows == 3)
            klaunch(ax_rows_kernel<8, 3, 1, KAHAN, WIDE, Epi>, nblk, 256, 0, s, g, x, y_det, z_det, epi);
        else if (sh.lpr == 8 && sh.rows == 2)
            klaunch(ax_rows_kernel<8, 2, 2, KAHAN, WIDE, Epi>, nblk, 256, 0, s, g, x, y_det, z_det, epi);
        else {
            set_error(ASD_ERR_ARG, "no row-bundle forward projector for this shape");
            return ASD_ERR_ARG;
        }
        ASD_LAUNCH_CHECK("ax_rows_kernel");
        return ASD_OK;
    }
    const int lpr = sh.lpr;
    long nrays = (long)g.na * g.nb;
    auto blocks = [&](int l) { return (unsigned)((nrays * l + 255) / 256); };
    switch (lpr) {
    case 1: klaunch(ax_kernel<1, KAHAN, WIDE, Epi>, blocks(1), 256, 0, s, g, x, y_det, z_det, epi); break;
    case 2: klaunch(ax_kernel<2, KAHAN, WIDE, Epi>, blocks(2), 256, 0, s, g, x, y_det, z_det, epi); break;
    case 4: klaunch(ax_kernel<4, KAHAN, WIDE, Epi>, blocks(4), 256, 0, s, g, x, y_det, z_det, epi); break;
    case 8: klaunch(ax_kernel<8, KAHAN, WIDE, Epi>, blocks(8), 256, 0, s, g, x, y_det, z_det, epi); break;
    case 16: klaunch(ax_kernel<16, KAHAN, WIDE, Epi>, blocks(16), 256, 0, s, g, x, y_det, z_det, epi); break;
    case 32: klaunch(ax_kernel<32, KAHAN, WIDE, Epi>, blocks(32), 256, 0, s, g, x, y_det, z_det, epi); break;
    case 64: klaunch(ax_kernel<64, KAHAN, WIDE, Epi>, blocks(64), 256, 0, s, g, x, y_det, z_det, epi); break;
    case 128: klaunch(ax_kernel<128, KAHAN, WIDE, Epi>, blocks(128), 256, 0, s, g, x, y_det, z_det, epi); break;
    default: klaunch(ax_kernel<256, KAHAN, WIDE, Epi>, blocks(256), 256, 0, s, g, x, y_det, z_det, epi); break;
    }
    ASD_LAUNCH_CHECK("ax_kernel");
    return ASD_OK;
}

// Forward projector launch for one view (shape from ax_shape): buffer reads
// where the volume allows them (vol_buffer_ok), 64-bit addressing otherwise.
template <bool KAHAN = false, class Epi>
int dispatch_ax(AxShape sh, const ViewGeom &g, const float *x, const float *y_det,
                const float *z_det, const Epi &epi, hipStream_t s)
{
    if (vol_buffer_ok((size_t)g.nx * g.ny * g.nzs, g.nx))
        return dispatch_ax_t<KAHAN, false>(sh, g, x, y_det, z_det, epi, s);
    return dispatch_ax_t<KAHAN, true>(sh, g, x, y_det, z_det, epi, s);
}

// Backprojector: one thread per voxel (x fastest) walks its detector
// footprint, ib outer and ia inner like the reference (cu:255-269), and sums
// l*y and l in that order.  The footprint's (ib, ia) pairs are taken in that
// order in chunks of ATX_CHUNK whose records and corrections are loaded
// together (clamped to the last pair) before any is used, so a voxel whose
// cut footprint holds a handful of pairs costs one memory round trip.  On
// views whose footprints are not cut, bins whose in-plane chord provably
// misses the voxel are skipped (their term is an exact zero).  Blocks are
// remapped so each XCD owns a contiguous slab of voxels and the rays they
// share stay in its L2.
constexpr int ATX_CHUNK = 4;

// Structured buffer loads (llvm.amdgcn.struct.ptr.buffer.load: the element
// index goes to the instruction's index operand, which the hardware scales by
// the descriptor's stride).
using u32x4 = __attribute__((ext_vector_type(4))) unsigned;
__device__ u32x4 struct_buffer_load_b128(__amdgpu_buffer_rsrc_t rsrc, unsigned index,
                                         unsigned voffset, unsigned soffset, int aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.v4i32");
__device__ float struct_buffer_load_f32(__amdgpu_buffer_rsrc_t rsrc, unsigned index,
                                        unsigned voffset, unsigned soffset, int aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.f32");

// Reads of a view's ray records and corrections by bin index (fewer than
// 2^27 bins): structured buffer resources with strides of 32 and 4 bytes, so
// the bin index is the load's operand as it is, with no address arithmetic.
// Pointer form otherwise.  Each resource holds the view's M records (the
// hardware checks a strided load's index against the record count); the
// walks only read bins of the view.
struct BinsBuf {
    __amdgpu_buffer_rsrc_t rr, ry;
    __device__ __forceinline__ BinsBuf(const RayRec *rt, const float *Y, unsigned M)
    {
        static_assert(sizeof(RayRec) == 32, "record stride");
        // dword3 0x00020000: 32-bit data format, the form untyped loads use;
        // both buffers end at bin M, so a read past it would return zeros
        rr = __builtin_amdgcn_make_buffer_rsrc((void *)rt, (short)32, (int)M, 0x00020000);
        ry = __builtin_amdgcn_make_buffer_rsrc((void *)Y, (short)4, (int)M, 0x00020000);
    }
    __device__ __forceinline__ RayRec rec(unsigned i) const
    {
        struct Q {
            u32x4 a, b;
        } q;
        q.a = struct_buffer_load_b128(rr, i, 0, 0, 0);
        q.b = struct_buffer_load_b128(rr, i, 16, 0, 0);
        return __builtin_bit_cast(RayRec, q);
    }
    __device__ __forceinline__ float y(unsigned i) const
    {
        return struct_buffer_load_f32(ry, i, 0, 0, 0);
    }
};

struct BinsPtr {
    const RayRec *__restrict__ rt;
    const float *__restrict__ Y;
    __host__ __device__ __forceinline__ RayRec rec(unsigned i) const { return rt[i]; }
    __host__ __device__ __forceinline__ float y(unsigned i) const { return Y[i]; }
};

// The footprint walk on a view whose rays are all general rays with exact
// face offsets (ViewGeom::general): the chord of every pair is chord3_sides,
// with no dispatch on the ray's mode, and the pairs past the footprint's end
// in the last load batch are evaluated and added as nothing (Acc::add_if)
// instead of branched around.  Same terms in the same order as the walk below.
template <bool KAHAN, int CHUNK, class Bins>
__host__ __device__ __forceinline__ void atx_voxel_walk_general(const ViewGeom &g, const Bins &B,
                                                       const Footprint &f, float px, float py,
                                                       float pz, Acc<KAHAN> &acc,
                                                       Acc<KAHAN> &accl)
{
#pragma clang fp contract(off)
    // a view of general rays has dz == 1 (view_general): half a voxel on
    // every axis, as a compile-time constant
    constexpr float hz = 0.5f;
    const int npairs = (f.a_hi - f.a_lo + 1) * (f.b_hi - f.b_lo + 1);
    int ia = f.a_lo;
    unsigned idx = mad24((unsigned)f.b_lo, (unsigned)g.na, (unsigned)f.a_lo);
    const unsigned jump = (unsigned)(g.na - (f.a_hi - f.a_lo));
    for (int k0 = 0; k0 < npairs; k0 += CHUNK) {
        RayRec R[CHUNK];
        float yv[CHUNK];
#pragma unroll
        for (int k = 0; k < CHUNK; k++) {
            R[k] = B.rec(idx);
            yv[k] = B.y(idx);
            // with one pair per batch the cursor past the last pair is never read
            if (CHUNK == 1 || k0 + k + 1 < npairs) {
                const bool wrap = ia == f.a_hi;
                ia = wrap ? f.a_lo : ia + 1;
                idx += wrap ? jump : 1u;
            }
        }
#pragma unroll
        for (int k = 0; k < CHUNK; k++) {
            const bool ok = k0 + k < npairs;
            const float l = chord3_sides(R[k], px, py, pz, hz);
            acc.add_if(ok, l * yv[k]);
            accl.add_if(ok, l);
        }
    }
}

// FAN: the kernel only ever sees fan views (the batch), whose rays are flat
// or take the a == 0 / b == 0 fallbacks: it leaves out the walk for views of
// general rays and tests flat rays first (find_l_3d_rec<true>), and with them
// the registers the general chord would hold.
template <bool KAHAN, int CHUNK, bool FAN, class Bins>
__device__ __forceinline__ void atx_voxel_walk_t(const ViewGeom &g, const Bins &B,
                                                 const Footprint &f, float px, float py,
                                                 float pz, float hz, Acc<KAHAN> &acc,
                                                 Acc<KAHAN> &accl)
{
#pragma clang fp contract(off)
    if (!FAN && g.general) {
        atx_voxel_walk_general<KAHAN, CHUNK>(g, B, f, px, py, pz, acc, accl);
        return;
    }
    const int npairs = (f.a_hi - f.a_lo + 1) * (f.b_hi - f.b_lo + 1);
    // cursor: the next pair in reference order, as its column ia and its bin
    // index ib * na + ia (a row wrap jumps to column a_lo of the next row)
    int ia = f.a_lo;
    unsigned idx = mad24((unsigned)f.b_lo, (unsigned)g.na, (unsigned)f.a_lo);
    const unsigned jump = (unsigned)(g.na - (f.a_hi - f.a_lo));
    for (int k0 = 0; k0 < npairs; k0 += CHUNK) {
        RayRec R[CHUNK];
        float yv[CHUNK];
#pragma unroll
        for (int k = 0; k < CHUNK; k++) {
            R[k] = B.rec(idx);
            yv[k] = B.y(idx);
            // advance unless this is the last pair (the rest of the chunk
            // re-reads it and is not used)
            if (k0 + k + 1 < npairs) {
                const bool wrap = ia == f.a_hi;
                ia = wrap ? f.a_lo : ia + 1;
                idx += wrap ? jump : 1u;
            }
        }
#pragma unroll
        for (int k = 0; k < CHUNK; k++) {
            if (k0 + k >= npairs || (!g.tight && !inplane_may_hit(R[k], px, py))) continue;
            float l = find_l_3d_rec<FAN>(R[k], px, py, pz, hz, g.dz);
            acc.add(l * yv[k]);
            accl.add(l);
        }
    }
}

// The footprint walk of one voxel (shared by atx_kernel, atx_sorted_kernel
// and the fallback of atx_pairs_kernel), CHUNK pairs per load batch.
template <bool KAHAN, int CHUNK = ATX_CHUNK, bool FAN = false>
__device__ __forceinline__ void atx_voxel_walk(const ViewGeom &g, const float *__restrict__ Y,
                                               const RayRec *__restrict__ rt, const Footprint &f,
                                               float px, float py, float pz, float hz,
                                               Acc<KAHAN> &acc, Acc<KAHAN> &accl)
{
    if (!(f.a_lo <= f.a_hi && f.b_lo <= f.b_hi)) return;
    const unsigned M = (unsigned)g.na * (unsigned)g.nb;  // < 2^31 (index_fits)
    if (M < (1u << 27))
        atx_voxel_walk_t<KAHAN, CHUNK, FAN>(g, BinsBuf(rt, Y, M), f, px, py, pz, hz, acc, accl);
    else
        atx_voxel_walk_t<KAHAN, CHUNK, FAN>(g, BinsPtr{rt, Y}, f, px, py, pz, hz, acc, accl);
}

template <class Epi, bool KAHAN = false, int CHUNK = ATX_CHUNK, bool FAN = false>
__global__ __launch_bounds__(256) void atx_kernel(ViewGeom g, const float *__restrict__ Y,
                                                  const RayRec *__restrict__ rt, Epi epi)
{
#pragma clang fp contract(off)
    const bool stop = epi.skip();
    const unsigned b = xcd_block(blockIdx.x, gridDim.x);
    const unsigned v = b * 256u + threadIdx.x;
    // the voxels of the z-window (the whole volume, or a slab's slices)
    const unsigned N = (unsigned)g.nx * (unsigned)g.ny * (unsigned)g.nzs;
    if (v >= N) return;
    const typename Epi::Pre pre = epi.load(v);
    // r < ny*nz < 2^24 and iz < nz (index_fits): 24-bit products
    const unsigned r = fastdiv(v, g.div_nx), iz = fastdiv(r, g.div_ny);
    const int ix = (int)(v - mad24(r, (unsigned)g.nx, 0u));
    const int iy = (int)(r - mad24(iz, (unsigned)g.ny, 0u));
    Footprint f = footprint(g, ix, iy, (int)iz + g.z0);
    const float sx = g.cos_phi * (-g.SO), sy = g.sin_phi * (-g.SO);
    const float px = sx - f.xc, py = sy - f.yc, pz = 0.0f - f.zc;
    Acc<KAHAN> acc, accl;
    atx_voxel_walk<KAHAN, CHUNK, FAN>(g, Y, rt, f, px, py, pz, g.dz / 2.0f, acc, accl);
    if (!stop) epi(v, acc.value(), accl.value(), pre);
}

// Backprojector with the work spread over (voxel, bin) pairs: a workgroup
// lists the pairs of its 256 voxels' cut footprints (each voxel's pairs in
// reference order), evaluates their lengths four per thread with all loads
// issued first, and parks the terms l*y and l in LDS; then every voxel adds
// its own terms in order.  The sums are the same FP32 additions in the same
// order as atx_kernel (bit-identical), but a wave no longer runs as many
// length evaluations as its largest footprint needs.  Blocks with more than
// PAIRS_MAX pairs, and views without the footprint cut, walk per voxel.
constexpr int PAIRS_MAX = 2048;
constexpr int PAIRS_PER_THREAD = 4;

template <class Epi, bool KAHAN = false>
__global__ __launch_bounds__(256) void atx_pairs_kernel(ViewGeom g, const float *__restrict__ Y,
                                                        const RayRec *__restrict__ rt, Epi epi)
{
#pragma clang fp contract(off)
    __shared__ float s_t[PAIRS_MAX], s_l[PAIRS_MAX];
    __shared__ unsigned s_bin[PAIRS_MAX];
    __shared__ unsigned char s_own[PAIRS_MAX];
    __shared__ float s_p[3][256];
    __shared__ int s_wave[4];
    const bool stop = epi.skip();
    const unsigned b = xcd_block(blockIdx.x, gridDim.x);
    const unsigned v = b * 256u + threadIdx.x;
    const unsigned N = (unsigned)g.nx * (unsigned)g.ny * (unsigned)g.nz;
    const bool valid = v < N;
    typename Epi::Pre pre{};
    if (valid) pre = epi.load(v);
    Footprint f;
    f.a_lo = f.b_lo = 0;
    f.a_hi = f.b_hi = -1;
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    if (valid) {
        const unsigned r = fastdiv(v, g.div_nx), iz = fastdiv(r, g.div_ny);
        f = footprint(g, (int)(v - mad24(r, (unsigned)g.nx, 0u)),
                      (int)(r - mad24(iz, (unsigned)g.ny, 0u)), (int)iz);
        px = g.cos_phi * (-g.SO) - f.xc;
        py = g.sin_phi * (-g.SO) - f.yc;
        pz = 0.0f - f.zc;
    }
    const int np = (f.a_lo <= f.a_hi && f.b_lo <= f.b_hi)
                       ? (f.a_hi - f.a_lo + 1) * (f.b_hi - f.b_lo + 1) : 0;
    // exclusive scan of np over the workgroup
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int incl = np;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(incl, d, 64);
        if (lane >= d) incl += t;
    }
    if (lane == 63) s_wave[w] = incl;
    __syncthreads();
    int base = 0, total = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        base += k < w ? s_wave[k] : 0;
        total += s_wave[k];
    }
    const int off = base + incl - np;
    const float hz = g.dz / 2.0f;
    Acc<KAHAN> acc, accl;
    if (!g.tight || total > PAIRS_MAX) {
        if (valid) atx_voxel_walk<KAHAN>(g, Y, rt, f, px, py, pz, hz, acc, accl);
    } else {
        // list this voxel's pairs (reference order) and its position
        s_p[0][threadIdx.x] = px;
        s_p[1][threadIdx.x] = py;
        s_p[2][threadIdx.x] = pz;
        int ia = f.a_lo, ib = f.b_lo;
        for (int k = 0; k < np; k++) {
            s_bin[off + k] = (unsigned)ib * (unsigned)g.na + (unsigned)ia;
            s_own[off + k] = (unsigned char)threadIdx.x;
            const bool wrap = ia == f.a_hi;
            ia = wrap ? f.a_lo : ia + 1;
            ib = wrap ? ib + 1 : ib;
        }
        __syncthreads();
        for (int p0 = threadIdx.x; p0 < total; p0 += 256 * PAIRS_PER_THREAD) {
            RayRec R[PAIRS_PER_THREAD];
            float yv[PAIRS_PER_THREAD];
#pragma unroll
            for (int j = 0; j < PAIRS_PER_THREAD; j++) {
                const int p = p0 + 256 * j;
                const unsigned idx = s_bin[p < total ? p : p0];
                R[j] = rt[idx];
                yv[j] = Y[idx];
            }
#pragma unroll
            for (int j = 0; j < PAIRS_PER_THREAD; j++) {
                const int p = p0 + 256 * j;
                if (p >= total) continue;
                const int o = s_own[p];
                const float l = find_l_3d_rec(R[j], s_p[0][o], s_p[1][o], s_p[2][o], hz, g.dz);
                s_t[p] = l * yv[j];
                s_l[p] = l;
            }
        }
        __syncthreads();
        for (int k = 0; k < np; k++) {
            acc.add(s_t[off + k]);
            accl.add(s_l[off + k]);
        }
    }
    if (valid && !stop) epi(v, acc.value(), accl.value(), pre);
}

// Backprojector with the voxels of a tile regrouped by footprint size.  A
// wave runs as many pair evaluations as its largest footprint needs; cut 3-D
// footprints hold 2 to 12 pairs, so a wave of neighbouring voxels runs about
// 1.5 times the pairs its voxels have.  Here a workgroup owns a TX x TY x TZ
// tile, computes every footprint, counting-sorts its voxels by pair
// count (one LDS atomic per voxel, largest first), and thread t then walks
// the voxel in sorted slot t, so each wave's voxels have similar footprints
// while the tile's rays stay few (the tile is compact on the detector).
// Each voxel's walk and sum are those of atx_kernel (bit-identical); only
// the lane a voxel runs on changes.  With TZ > 1 the in-plane part of the
// footprints (footprint_col, the costlier part) is computed once per column
// of the tile by its first TX x TY threads and shared through LDS.
constexpr int SORT_CLASSES = 32;  // pair counts 0..30, and 31 for "31 or more"

template <int TX, int TY, int TZ, class Epi, bool KAHAN = false, int CHUNK = 2>
__global__ __launch_bounds__(TX *TY *TZ) void atx_sorted_kernel(ViewGeom g,
                                                                 const float *__restrict__ Y,
                                                                 const RayRec *__restrict__ rt,
                                                                 Epi epi)
{
#pragma clang fp contract(off)
    constexpr int T = TX * TY * TZ;
    __shared__ int s_cnt[SORT_CLASSES], s_off[SORT_CLASSES];
    __shared__ unsigned s_v[T];
    __shared__ int s_f[4][T];
    __shared__ float s_p[3][T];
    __shared__ FootCol s_col[TZ > 1 ? TX * TY : 1];
    const bool stop = epi.skip();
    const TileDims td = tile_dims(g, TX, TY, TZ);
    const unsigned t = xcd_block(blockIdx.x, gridDim.x);
    const int tile_x = (int)(t % td.ntx), tile_y = (int)((t / td.ntx) % td.nty),
              tile_z = (int)(t / ((unsigned)td.ntx * td.nty));
    const int ix = tile_x * TX + (int)(threadIdx.x % TX),
              iy = tile_y * TY + (int)((threadIdx.x / TX) % TY),
              iz = tile_z * TZ + (int)(threadIdx.x / (TX * TY));
    const bool valid = ix < g.nx && iy < g.ny && iz < g.nz;
    if constexpr (TZ > 1) {
        if ((int)threadIdx.x < TX * TY && ix < g.nx && iy < g.ny)
            s_col[threadIdx.x] = footprint_col(g, ix, iy);
        __syncthreads();
    }
    Footprint f;
    f.a_lo = f.b_lo = 0;
    f.a_hi = f.b_hi = -1;
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    if (valid) {
        if constexpr (TZ > 1)
            f = footprint_z(g, s_col[threadIdx.x % (TX * TY)], iz);
        else
            f = footprint(g, ix, iy, iz);
        px = g.cos_phi * (-g.SO) - f.xc;
        py = g.sin_phi * (-g.SO) - f.yc;
        pz = 0.0f - f.zc;
    }
    const int np = (f.a_lo <= f.a_hi && f.b_lo <= f.b_hi)
                       ? (f.a_hi - f.a_lo + 1) * (f.b_hi - f.b_lo + 1) : 0;
    const int cls = np < SORT_CLASSES - 1 ? np : SORT_CLASSES - 1;
    if (threadIdx.x < SORT_CLASSES) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const int pos = atomicAdd(&s_cnt[cls], 1);
    __syncthreads();
    if (threadIdx.x < 64) {
        // exclusive offsets of the classes, largest class first
        const int lane = threadIdx.x, c = SORT_CLASSES - 1 - lane;
        const int n = lane < SORT_CLASSES ? s_cnt[c] : 0;
        int incl = n;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int up = __shfl_up(incl, d, 64);
            if (lane >= d) incl += up;
        }
        if (lane < SORT_CLASSES) s_off[c] = incl - n;
    }
    __syncthreads();
    const int slot = s_off[cls] + pos;
    s_v[slot] = valid ? (unsigned)(((size_t)iz * g.ny + iy) * g.nx + ix) : ~0u;
    s_f[0][slot] = f.a_lo;
    s_f[1][slot] = f.a_hi;
    s_f[2][slot] = f.b_lo;
    s_f[3][slot] = f.b_hi;
    s_p[0][slot] = px;
    s_p[1][slot] = py;
    s_p[2][slot] = pz;
    __syncthreads();
    const unsigned vv = s_v[threadIdx.x];
    if (vv == ~0u) return;
    const typename Epi::Pre pre = epi.load(vv);
    Footprint fs;
    fs.a_lo = s_f[0][threadIdx.x];
    fs.a_hi = s_f[1][threadIdx.x];
    fs.b_lo = s_f[2][threadIdx.x];
    fs.b_hi = s_f[3][threadIdx.x];
    Acc<KAHAN> acc, accl;
    atx_voxel_walk<KAHAN, CHUNK>(g, Y, rt, fs, s_p[0][threadIdx.x], s_p[1][threadIdx.x],
                                 s_p[2][threadIdx.x], g.dz / 2.0f, acc, accl);
    if (!stop) epi(vv, acc.value(), accl.value(), pre);
}

// atx_sorted_kernel for views of general rays (ViewGeom::general) with fewer
// than 2048 bins per detector row and column.  A workgroup of 256 threads
// owns a tile of 512 voxels, two per thread:
//  - the columns' footprints (footprint_col) and the tile's counters and
//    barriers are shared by twice the voxels;
//  - the 512 voxels are regrouped by pair count, largest first, and thread t
//    walks sorted slots t and 511 - t one after the other, so each wave's
//    lanes hold footprints of similar size in both walks, and the waves'
//    two walks together are of similar length;
//  - a slot carries two words, the footprint's first bin (a_lo, b_lo) and its
//    widths with the voxel's index in the tile, instead of eight; the thread
//    that takes the slot recomputes the voxel centre from the index with the
//    expressions footprint_col / footprint_z use (same bits);
//  - the tile index is split by multiply-high division (FastDiv) on the
//    scalar unit, and the class counters are cleared before the column
//    barrier;
//  - each walk takes one pair per load batch, evaluated without a mode
//    dispatch (atx_voxel_walk_general).
// Each voxel's terms and their order are those of atx_sorted_kernel.
// Tile t (from xcd_block, which gives each XCD a contiguous range of tiles) of
// a backprojector launch whose tiles are ordered x fastest, then y, then
// layer (TZ slices): XCD x takes layers x, x + 8, x + 16, ... instead of one
// contiguous band of layers.  Voxels near the top and
// bottom of the volume, where the cone leaves the detector, have fewer pairs
// (at 512^3 the outer eighths hold 75-78% of the mean work, the others
// 107-108%), and the launch ends with the XCD holding the most work; spread
// over the whole height every XCD holds about the same.  The map is a
// bijection on [0, per_layer * nlayers) (checked on the CPU); layer counts
// that are not a multiple of 8 (or below 16) keep xcd_block's order.  Blocks
// of 4 layers per XCD turn (16x16x2 tiles at 512^3) were 1.6% slower than
// single layers.
__host__ __device__ __forceinline__ unsigned atx_tile_z_interleave(unsigned t, unsigned per_layer,
                                                                   unsigned nlayers)
{
    const unsigned B = nlayers;  // blocks of one layer
    if (B % 8u != 0 || B < 16u) return t;
    const unsigned k = nlayers / B, q = B / 8u;
    const unsigned layer = t / per_layer, in = t - layer * per_layer;
    const unsigned j = layer / k, lk = layer - j * k;
    const unsigned x = j / q, i = j - x * q;
    return ((i * 8u + x) * k + lk) * per_layer + in;
}

template <int TX, int TY, int TZ, class Epi, bool KAHAN = false>
__global__ __launch_bounds__(256) void atx_sorted_general_kernel(
    ViewGeom g, const float *__restrict__ Y, const RayRec *__restrict__ rt, Epi epi,
    FastDiv div_ntx, FastDiv div_nty)
{
#pragma clang fp contract(off)
    constexpr int T = TX * TY * TZ;
    static_assert(T == 512 && TX * TY <= 256 && 256 % (TX * TY) == 0,
                  "two voxels per thread, tile index in 9 bits");
    constexpr int ZH = 256 / (TX * TY);  // slices between a thread's two voxels
    __shared__ int s_cnt[SORT_CLASSES], s_off[SORT_CLASSES];
    __shared__ unsigned s_w0[T], s_w1[T];
    __shared__ FootCol s_col[TX * TY];
    const bool stop = epi.skip();
    const unsigned ntx = (unsigned)(g.nx + TX - 1) / TX, nty = (unsigned)(g.ny + TY - 1) / TY;
    const unsigned t = atx_tile_z_interleave(xcd_block(blockIdx.x, gridDim.x), ntx * nty,
                                             (unsigned)(g.nz + TZ - 1) / TZ);
    const unsigned tq = fastdiv(t, div_ntx), tile_x = t - tq * ntx;
    const unsigned tile_z = fastdiv(tq, div_nty), tile_y = tq - tile_z * nty;
    const int ix = (int)tile_x * TX + (int)(threadIdx.x % TX),
              iy = (int)tile_y * TY + (int)((threadIdx.x / TX) % TY),
              lz = (int)(threadIdx.x / (TX * TY));
    if (threadIdx.x < SORT_CLASSES) s_cnt[threadIdx.x] = 0;
    if ((int)threadIdx.x < TX * TY && ix < g.nx && iy < g.ny)
        s_col[threadIdx.x] = footprint_col(g, ix, iy);
    __syncthreads();
    unsigned w0[2], w1[2];
    int cls[2], pos[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int iz = (int)tile_z * TZ + lz + h * ZH;
        const bool valid = ix < g.nx && iy < g.ny && iz < g.nz;
        Footprint f;
        f.a_lo = f.b_lo = 0;
        f.a_hi = f.b_hi = -1;
        if (valid) f = footprint_z(g, s_col[threadIdx.x % (TX * TY)], iz);
        const bool nonempty = f.a_lo <= f.a_hi && f.b_lo <= f.b_hi;
        const unsigned wa = nonempty ? (unsigned)(f.a_hi - f.a_lo + 1) : 0u,
                       wb = nonempty ? (unsigned)(f.b_hi - f.b_lo + 1) : 0u;
        const int np = (int)(wa * wb);
        cls[h] = np < SORT_CLASSES - 1 ? np : SORT_CLASSES - 1;
        pos[h] = atomicAdd(&s_cnt[cls[h]], 1);
        // voxels outside the volume: first bin word all ones (a_lo < 2047)
        w0[h] = valid ? (unsigned)f.a_lo | (unsigned)f.b_lo << 16 : ~0u;
        w1[h] = (threadIdx.x + h * 256u) | wa << 9 | wb << 20;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        // exclusive offsets of the classes, largest class first
        const int lane = threadIdx.x, c = SORT_CLASSES - 1 - lane;
        const int n = lane < SORT_CLASSES ? s_cnt[c] : 0;
        int incl = n;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int up = __shfl_up(incl, d, 64);
            if (lane >= d) incl += up;
        }
        if (lane < SORT_CLASSES) s_off[c] = incl - n;
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int slot = s_off[cls[h]] + pos[h];
        s_w0[slot] = w0[h];
        s_w1[slot] = w1[h];
    }
    __syncthreads();
    const unsigned M = (unsigned)g.na * (unsigned)g.nb;  // < 2^22
    const BinsBuf B(rt, Y, M);
#pragma unroll 1
    for (int h = 0; h < 2; h++) {
        // the second walk takes the sorted slots in reverse, so wave 0 pairs
        // the largest footprints of the first half with the smallest of the
        // second: the four waves of a workgroup finish closer together
        const unsigned sl = h ? 511u - threadIdx.x : threadIdx.x;
        const unsigned a0 = s_w0[sl], a1 = s_w1[sl];
        if (a0 == ~0u) continue;
        const int me = (int)(a1 & 511u);
        const int jx = (int)tile_x * TX + me % TX, jy = (int)tile_y * TY + (me / TX) % TY,
                  jz = (int)tile_z * TZ + me / (TX * TY);
        const unsigned vv = (unsigned)(((size_t)jz * g.ny + jy) * g.nx + jx);
        const typename Epi::Pre pre = epi.load(vv);
        Footprint fs;
        fs.a_lo = (int)(a0 & 0xffffu);
        fs.b_lo = (int)(a0 >> 16);
        fs.a_hi = fs.a_lo + (int)((a1 >> 9) & 2047u) - 1;
        fs.b_hi = fs.b_lo + (int)(a1 >> 20) - 1;
        // the voxel centre as footprint_col and footprint_z form it
        const float px = g.cos_phi * (-g.SO) - voxel_xc(g, jx),
                    py = g.sin_phi * (-g.SO) - voxel_yc(g, jy), pz = 0.0f - voxel_zc(g, jz);
        Acc<KAHAN> acc, accl;
        if (fs.a_lo <= fs.a_hi && fs.b_lo <= fs.b_hi)
            atx_voxel_walk_general<KAHAN, 1>(g, B, fs, px, py, pz, acc, accl);
        if (!stop) epi(vv, acc.value(), accl.value(), pre);
    }
}

// Whether atx_sorted_general_kernel may take a view: general rays, footprint
// bounds below 2048 (16-bit first bins, 11- and 12-bit widths in the slot
// words), and the whole volume (no z-window).
__host__ __device__ __forceinline__ bool atx_general_ok(const ViewGeom &g)
{
    return g.general && g.na < 2048 && g.nb < 2048 && g.z0 == 0 && g.nzs == g.nz;
}

// Backprojector on voxel tiles (TX x TY x TZ = 256 voxels per workgroup):
// the slope records and correction values of the tile's footprint are
// staged in LDS once (coalesced), then every voxel walks its own footprint,
// ib outer and ia inner like the reference, from LDS.  Same arithmetic and
// summation order as atx_kernel, so the result is bit-identical to it.
template <int TX, int TY, int TZ, bool KAHAN, class Epi>
__global__ __launch_bounds__(256) void atx_tile_plain_kernel(ViewGeom g,
                                                             const float *__restrict__ Y,
                                                             const RayRec *__restrict__ rt,
                                                             Epi epi)
{
#pragma clang fp contract(off)
    static_assert(TX * TY * TZ == 256, "one voxel per thread");
    __shared__ RayRec s_rec[TILE_MAXR];
    __shared__ float s_y[TILE_MAXR];
    __shared__ int s_rng[4];
    const bool stop = epi.skip();
    const TileDims td = tile_dims(g, TX, TY, TZ);
    const unsigned ntiles = (unsigned)td.ntx * td.nty * td.ntz;
    const unsigned t = xcd_block(blockIdx.x, gridDim.x);
    if (t >= ntiles) return;
    const int tile_x = (int)(t % td.ntx), tile_y = (int)((t / td.ntx) % td.nty),
              tile_z = (int)(t / ((unsigned)td.ntx * td.nty));
    const int lx = threadIdx.x % TX, ly = (threadIdx.x / TX) % TY, lz = threadIdx.x / (TX * TY);
    const int ix = tile_x * TX + lx, iy = tile_y * TY + ly, iz = tile_z * TZ + lz;
    const bool valid = ix < g.nx && iy < g.ny && iz < g.nz;
    const long vox = (long)((size_t)iz * g.ny + iy) * g.nx + ix;
    typename Epi::Pre pre{};
    if (valid) pre = epi.load(vox);
    Footprint f;
    if (valid) f = footprint(g, ix, iy, iz);
    tile_union(valid && f.a_lo <= f.a_hi && f.b_lo <= f.b_hi, f, s_rng);
    const int A0 = s_rng[0], A1 = s_rng[1], B0 = s_rng[2], B1 = s_rng[3];
    const float sx = g.cos_phi * (-g.SO), sy = g.sin_phi * (-g.SO);
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    if (valid) {
        px = sx - f.xc;
        py = sy - f.yc;
        pz = 0.0f - f.zc;
    }
    const float hz = g.dz / 2.0f;
    Acc<KAHAN> acc, accl;
    if (A0 <= A1 && B0 <= B1) {
        const int nA = A1 - A0 + 1;
        if (nA > TILE_MAXR) {
            // footprint too wide for LDS: read the records directly
            if (valid)
                for (int ib = f.b_lo; ib <= f.b_hi; ib++)
                    for (int ia = f.a_lo; ia <= f.a_hi; ia++) {
                        const long k = (long)ib * g.na + ia;
                        const RayRec R = rt[k];
                        if (!inplane_may_hit(R, px, py)) continue;
                        float l = find_l_3d_rec(R, px, py, pz, hz, g.dz);
                        acc.add(l * Y[k]);
                        accl.add(l);
                    }
        } else {
            const int RB = TILE_MAXR / nA;  // detector rows per LDS chunk
            for (int b0 = B0; b0 <= B1; b0 += RB) {
                const int b1 = b0 + RB - 1 < B1 ? b0 + RB - 1 : B1;
                const int n = nA * (b1 - b0 + 1);
                __syncthreads();
                for (int k = threadIdx.x; k < n; k += 256) {
                    const long gk = (long)(b0 + k / nA) * g.na + A0 + k % nA;
                    s_rec[k] = rt[gk];
                    s_y[k] = Y[gk];
                }
                __syncthreads();
                if (valid) {
                    const int blo = f.b_lo > b0 ? f.b_lo : b0, bhi = f.b_hi < b1 ? f.b_hi : b1;
                    for (int ib = blo; ib <= bhi; ib++) {
                        const int row = (ib - b0) * nA - A0;
                        for (int ia = f.a_lo; ia <= f.a_hi; ia++) {
                            const RayRec R = s_rec[row + ia];
                            if (!g.tight && !inplane_may_hit(R, px, py)) continue;
                            float l = find_l_3d_rec(R, px, py, pz, hz, g.dz);
                            acc.add(l * s_y[row + ia]);
                            accl.add(l);
                        }
                    }
                }
            }
        }
    }
    if (valid && !stop) epi(vox, acc.value(), accl.value(), pre);
}

// Backprojector variants (all bit-identical; the plan's autotuner picks one).
enum : int {
    ATX_DIRECT = 0,   // one thread per voxel, records from L1/L2
    ATX_T16x16x1 = 1, // LDS-staged 16x16 slices
    ATX_T8x8x4 = 2,   // LDS-staged 8x8x4 bricks
    ATX_T16x4x4 = 3,  // LDS-staged 16x4x4 bricks
    ATX_PAIRS = 4,    // (voxel, bin) pairs spread over the workgroup
    ATX_DIRECT2 = 5,  // as ATX_DIRECT, two pairs per load batch
    ATX_DIRECT1 = 6,  // as ATX_DIRECT, one pair at a time
    ATX_SORTED16 = 7, // two-pair walks, 16x16x1 tiles regrouped by footprint size
    ATX_SORTED8 = 8,  // the same on 8x8x4 tiles
    ATX_SORTED8Z = 9, // the same on 8x4x8 tiles
    ATX_NVARIANT = 10,
};

template <int TX, int TY, int TZ, bool KAHAN, class Epi>
int launch_atx_tile_t(const ViewGeom &g, const float *Y, const RayRec *rt, const Epi &epi,
                      hipStream_t s)
{
    const TileDims td = tile_dims(g, TX, TY, TZ);
    klaunch(atx_tile_plain_kernel<TX, TY, TZ, KAHAN, Epi>, (unsigned)td.ntx * td.nty * td.ntz, 256,
            0, s, g, Y, rt, epi);
    ASD_LAUNCH_CHECK("atx_tile_plain_kernel");
    return ASD_OK;
}

template <bool KAHAN, class Epi>
int dispatch_atx_variant(int variant, const ViewGeom &g, const float *Y, const RayRec *rt,
                         const Epi &epi, hipStream_t s)
{
    switch (variant) {
    case ATX_T16x16x1: return launch_atx_tile_t<16, 16, 1, KAHAN>(g, Y, rt, epi, s);
    case ATX_T8x8x4: return launch_atx_tile_t<8, 8, 4, KAHAN>(g, Y, rt, epi, s);
    case ATX_T16x4x4: return launch_atx_tile_t<16, 4, 4, KAHAN>(g, Y, rt, epi, s);
    case ATX_DIRECT2: {
        const long N = (long)g.nx * g.ny * g.nzs;
        // one-slice views: the walk compiled for fan rays (find_l_3d_rec<true>)
        if (g.nz == 1)
            klaunch(atx_kernel<Epi, KAHAN, 2, true>, (unsigned)((N + 255) / 256), 256, 0, s, g, Y,
                    rt, epi);
        else
            klaunch(atx_kernel<Epi, KAHAN, 2>, (unsigned)((N + 255) / 256), 256, 0, s, g, Y, rt,
                    epi);
        ASD_LAUNCH_CHECK("atx_kernel");
        return ASD_OK;
    }
    case ATX_SORTED16:
    case ATX_SORTED8: {
        const bool t16 = variant == ATX_SORTED16;
        if (atx_general_ok(g)) {
            // the same column footprints over twice the slices
            const TileDims tg = t16 ? tile_dims(g, 16, 16, 2) : tile_dims(g, 8, 8, 8);
            const unsigned ng = (unsigned)tg.ntx * tg.nty * tg.ntz;
            const FastDiv dx = make_fastdiv((unsigned)tg.ntx), dy = make_fastdiv((unsigned)tg.nty);
            if (t16)
                klaunch(atx_sorted_general_kernel<16, 16, 2, Epi, KAHAN>, ng, 256, 0, s, g, Y, rt,
                        epi, dx, dy);
            else
                klaunch(atx_sorted_general_kernel<8, 8, 8, Epi, KAHAN>, ng, 256, 0, s, g, Y, rt,
                        epi, dx, dy);
            ASD_LAUNCH_CHECK("atx_sorted_general_kernel");
            return ASD_OK;
        }
        const TileDims td = t16 ? tile_dims(g, 16, 16, 1) : tile_dims(g, 8, 8, 4);
        const unsigned nt = (unsigned)td.ntx * td.nty * td.ntz;
        if (t16)
            klaunch(atx_sorted_kernel<16, 16, 1, Epi, KAHAN, 2>, nt, 256, 0, s, g, Y, rt, epi);
        else
            klaunch(atx_sorted_kernel<8, 8, 4, Epi, KAHAN, 2>, nt, 256, 0, s, g, Y, rt, epi);
        ASD_LAUNCH_CHECK("atx_sorted_kernel");
        return ASD_OK;
    }
    case ATX_SORTED8Z: {
        const TileDims td = tile_dims(g, 8, 4, 8);
        if (atx_general_ok(g)) {
            const TileDims tg = tile_dims(g, 8, 4, 16);
            klaunch(atx_sorted_general_kernel<8, 4, 16, Epi, KAHAN>,
                    (unsigned)tg.ntx * tg.nty * tg.ntz, 256, 0, s, g, Y, rt, epi,
                    make_fastdiv((unsigned)tg.ntx), make_fastdiv((unsigned)tg.nty));
            ASD_LAUNCH_CHECK("atx_sorted_general_kernel");
            return ASD_OK;
        }
        klaunch(atx_sorted_kernel<8, 4, 8, Epi, KAHAN, 2>, (unsigned)td.ntx * td.nty * td.ntz, 256,
                0, s, g, Y, rt, epi);
        ASD_LAUNCH_CHECK("atx_sorted_kernel");
        return ASD_OK;
    }
    case ATX_DIRECT1: {
        const long N = (long)g.nx * g.ny * g.nzs;
        if (g.nz == 1)
            klaunch(atx_kernel<Epi, KAHAN, 1, true>, (unsigned)((N + 255) / 256), 256, 0, s, g, Y,
                    rt, epi);
        else
            klaunch(atx_kernel<Epi, KAHAN, 1>, (unsigned)((N + 255) / 256), 256, 0, s, g, Y, rt,
                    epi);
        ASD_LAUNCH_CHECK("atx_kernel");
        return ASD_OK;
    }
    case ATX_PAIRS: {
        const long N = (long)g.nx * g.ny * g.nz;
        klaunch(atx_pairs_kernel<Epi, KAHAN>, (unsigned)((N + 255) / 256), 256, 0, s, g, Y, rt, epi);
        ASD_LAUNCH_CHECK("atx_pairs_kernel");
        return ASD_OK;
    }
    default: {
        const long N = (long)g.nx * g.ny * g.nzs;
        if (g.nz == 1)
            klaunch(atx_kernel<Epi, KAHAN, ATX_CHUNK, true>, (unsigned)((N + 255) / 256), 256, 0,
                    s, g, Y, rt, epi);
        else
            klaunch(atx_kernel<Epi, KAHAN>, (unsigned)((N + 255) / 256), 256, 0, s, g, Y, rt, epi);
        ASD_LAUNCH_CHECK("atx_kernel");
        return ASD_OK;
    }
    }
}

// Precision-variant backprojector (tile-local FP64 ray setup, see
// projector_tile.hpp).
template <bool KAHAN, class Epi>
int dispatch_atx_prec(const ViewGeom &g, const float *Y, const float *y_det, const float *z_det,
                      const RayRec *rt, const Epi &epi, hipStream_t s)
{
    if (g.nz == 1) {
        const TileDims td = tile_dims(g, 16, 16, 1);
        klaunch(atx_tile_kernel<16, 16, 1, KAHAN, Epi>, (unsigned)td.ntx * td.nty * td.ntz, 256, 0,
                s, g, Y, y_det, z_det, rt, epi);
    } else {
        const TileDims td = tile_dims(g, 8, 8, 4);
        klaunch(atx_tile_kernel<8, 8, 4, KAHAN, Epi>, (unsigned)td.ntx * td.nty * td.ntz, 256, 0, s,
                g, Y, y_det, z_det, rt, epi);
    }
    ASD_LAUNCH_CHECK("atx_tile_kernel");
    return ASD_OK;
}

}  // namespace asd
