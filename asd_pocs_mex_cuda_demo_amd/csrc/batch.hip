// Batched ASD-POCS: many independent reconstructions of uneven size and view
// count advance together, one kernel launch per stage for all of them.
//
// Cases are sorted by view count (descending), so at view step j the cases
// still sweeping are a prefix of that order.  Each stage walks a flat index
// space over its cases:
//   ax_batch    rays of view j of every active case (LPR lanes per ray)
//   atx_batch   256-voxel blocks of every active case (blocks never straddle cases)
//   tv / post   256-voxel blocks of every case, with per-case FP64 reductions
//               (the last block of a case folds that case's partials)
// Per-case alpha, beta, k and stop flag live in device memory, so an outer
// iteration of the whole batch is one fixed kernel sequence, captured as a
// hipGraph.  The arithmetic of every case is the single-case driver's.
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "projector_kernels.hpp"
#include "reduce_device.hpp"
#include "sino_codec.hpp"
#include "tv_stencil.hpp"

namespace asd {
namespace {

constexpr int B_HIST = 4096;  // RMSE entries kept per case

struct BCtl {
    double alpha, beta;
    double dp, dd2, dg, d2, nrm2;
    int k, stopped;
};

struct BRules {
    double alpha_red, beta_red, r_max, epsilon, stopRelChg;
    int ng;       // TV sub-iterations of this case
    int three_d;  // 3-D stencil
};

// One case at one view step.
struct BItem {
    ViewGeom g;
    float *m;
    const void *P;         // this view of the case's sinogram (see sino_codec.hpp)
    double lo, step;       // 16-bit fixed-point scale of this view
    float *cor;
    double *res2;
    RayRec *rt;
    const float *ydet, *zdet;
    BCtl *ctl;
};

// One case for the whole-volume stages.
struct BCase {
    int nx, ny, nz;
    int M;
    size_t voff, roff;
    BRules rules;
    int has_I0;
};

__device__ __forceinline__ int find_seg(const int *pref, int n, int x)
{
    // largest c < n with pref[c] <= x (pref ascending, pref[0] = 0)
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (pref[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <class Codec> __device__ __forceinline__ Codec item_codec(const BItem &it);
template <> __device__ __forceinline__ SinoF32 item_codec<SinoF32>(const BItem &it)
{
    return SinoF32{(const float *)it.P};
}
template <> __device__ __forceinline__ SinoU16 item_codec<SinoU16>(const BItem &it)
{
    return SinoU16{(const uint16_t *)it.P, it.lo, it.step};
}
template <> __device__ __forceinline__ SinoF16 item_codec<SinoF16>(const BItem &it)
{
    return SinoF16{(const __half *)it.P};
}

// Rays of one view step of every active case.  Each case's rays fill whole
// blocks (abpref: block prefix over the active cases), so the case, its work
// item and its stop flag are block-uniform (scalar loads).  The detector
// coordinates are the gateway grid computed in place; the ray tables were
// built once per case and view.
template <int LPR, bool KAHAN, class PT>
__global__ __launch_bounds__(256) void ax_batch_kernel(const BItem *__restrict__ items,
                                                       const int *__restrict__ abpref, int ncase)
{
    constexpr int RPB = 256 / LPR;
    const int lane = threadIdx.x % LPR;
    if ((int)blockIdx.x >= abpref[ncase]) return;
    const int c = find_seg(abpref, ncase, (int)blockIdx.x);
    const BItem &it = items[c];
    const ViewGeom &g = it.g;
    const long id = (long)((int)blockIdx.x - abpref[c]) * RPB + threadIdx.x / LPR;
    const bool active = id < (long)g.na * g.nb && !it.ctl->stopped;
    double pv = 0.0, r2 = 0.0;
    if (active && lane == 0) {  // epilogue operands, loaded before the march
        pv = item_codec<PT>(it).at(id);
        r2 = it.res2[id];
    }
    Acc<KAHAN> sa, sl;
    if (active) {
        const unsigned ub = fastdiv((unsigned)id, g.div_na);
        const int ia = (int)((unsigned)id - mad24(ub, (unsigned)g.na, 0u)), ib = (int)ub;
        AxRay r = ax_ray_setup(g, det_coord(ia, g.na, g.dy_det), det_coord(ib, g.nb, g.dz_det));
        // every case passes vol_buffer_ok (asd_batch_create)
        const VolRead<false> V(it.m, (size_t)g.nx * g.ny * g.nz);
        const SlabAddr A = slab_addr(r, g.nx, VolRead<false>::unit);
        int nu = r.xdriven ? g.nx : g.ny;
        auto march = [&](auto ft, auto flat) {
            for (int iu0 = lane; iu0 < nu; iu0 += 4 * LPR)
                ax_slabs<4, KAHAN, decltype(ft)::value, VolRead<false>, decltype(flat)::value>(
                    r, A, V, iu0, LPR, nu, g.ny, sa, sl);
        };
        // per-case constants (every ray of a 2-D case is flat): one branch
        // for the march
        if (r.fast_t && r.flat)
            march(std::true_type{}, std::true_type{});
        else if (r.fast_t)
            march(std::true_type{}, std::false_type{});
        else
            march(std::false_type{}, std::false_type{});
    }
    float acc, accl;
    if constexpr (LPR > 1 && KAHAN) {
        const double a = group_sum<LPR>((double)sa.s - (double)sa.c);
        const double b = group_sum<LPR>((double)sl.s - (double)sl.c);
        acc = (float)a;
        accl = (float)b;
    } else {
        acc = sa.value();
        accl = sl.value();
        if constexpr (LPR > 1) {
            acc = group_sum<LPR>(acc);
            accl = group_sum<LPR>(accl);
        }
    }
    if (!active || lane != 0) return;
    {
#pragma clang fp contract(off)
        float proj = acc * g.scale, sumr = accl * g.scale;
        double r = (sumr == 0.0f) ? 1.0 : (double)sumr;
        double c0 = pv - (double)proj;
        it.res2[id] = r2 + c0 * c0;
        it.cor[id] = (float)(c0 / r);
    }
}

// CHUNK: footprint pairs per load batch (the batch's autotuner picks it; every
// choice gives the same bits).  dry: compute everything but store nothing
// (the autotuner's timing runs).
template <bool KAHAN, int CHUNK>
__global__ __launch_bounds__(256) void atx_batch_kernel(const BItem *__restrict__ items,
                                                        const int *__restrict__ bpref, int ncase,
                                                        int dry)
{
#pragma clang fp contract(off)
    const int b = blockIdx.x;
    if (b >= bpref[ncase]) return;
    const int c = find_seg(bpref, ncase, b);
    const BItem &it = items[c];
    if (it.ctl->stopped) return;
    const ViewGeom &g = it.g;
    const long v = (long)(b - bpref[c]) * 256 + threadIdx.x;
    const long N = (long)g.nx * g.ny * g.nz;
    if (v >= N) return;
    const float m0 = it.m[v];
    const unsigned r = fastdiv((unsigned)v, g.div_nx), uz = fastdiv(r, g.div_ny);
    const int ix = (int)((unsigned)v - mad24(r, (unsigned)g.nx, 0u));
    const int iy = (int)(r - mad24(uz, (unsigned)g.ny, 0u)), iz = (int)uz;
    Footprint f = footprint(g, ix, iy, iz);
    const float sx = g.cos_phi * (-g.SO), sy = g.sin_phi * (-g.SO);
    const float px = sx - f.xc, py = sy - f.yc, pz = 0.0f - f.zc;
    Acc<KAHAN> acc, accl;
    // fan views only
    atx_voxel_walk<KAHAN, CHUNK, true>(g, it.cor, it.rt, f, px, py, pz, g.dz / 2.0f, acc, accl);
    float bp = acc.value() * g.scale, sc = accl.value() * g.scale;
    double cc = (sc == 0.0f) ? 1.0 : (double)sc;
    double nm = (double)m0 + it.ctl->beta * ((double)bp / cc);
    if (!dry) it.m[v] = nm < 0 ? 0.0f : (float)nm;
}

// The backprojector stage with `chunk` pairs per load batch (1, 2 or 4).
template <bool KAHAN>
void launch_atx_batch(int chunk, unsigned blocks, hipStream_t s, const BItem *items,
                      const int *bp, int A, int dry)
{
    if (chunk == 1)
        klaunch(atx_batch_kernel<KAHAN, 1>, blocks, 256, 0, s, items, bp, A, dry);
    else if (chunk == 4)
        klaunch(atx_batch_kernel<KAHAN, 4>, blocks, 256, 0, s, items, bp, A, dry);
    else
        klaunch(atx_batch_kernel<KAHAN, 2>, blocks, 256, 0, s, items, bp, A, dry);
}

// Per-case folds: block c of a <<<ncase>>> launch sums rows of the partials
// written by the blocks [bpref[c], bpref[c+1]) of the preceding stage.
template <int K>
__device__ __forceinline__ void fold_case(const double *partials, int nblocks_all, int first,
                                          int nblk, double (&tot)[K])
{
    __shared__ double res[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        double t = thread_row_sum(partials + (size_t)k * nblocks_all + first, nblk);
        t = block_sum(t);
        if (threadIdx.x == 0) res[k] = t;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) tot[k] = res[k];
}

// One TV sub-iteration `it` of every case, the step of sub-iteration it-1
// fused into gradient it as in the single-case descent (tv.hip):
//   it == 0          gradient of the case's image, FP64 partials of ||v||^2
//   0 < it < ng      fold the case's previous partials (in every block, fixed
//                    order) into s = alpha * d / ||v||, apply that step to
//                    every value read, store this voxel's stepped value and
//                    the new gradient
//   it == ng         the last step into the image, with positivity
// Blocks never straddle cases, so the case is block-uniform.
__global__ __launch_bounds__(256) void tv_sub_batch(const BCase *__restrict__ cs,
                                                    const int *__restrict__ bpref, int ncase,
                                                    float *__restrict__ u_all, float *__restrict__ v_all,
                                                    float *__restrict__ u2_all,
                                                    float *__restrict__ v2_all, const BCtl *ctl,
                                                    double *partials, int it)
{
#pragma clang fp contract(off)
    const int b = blockIdx.x;
    const int c = find_seg(bpref, ncase, b);
    const BCase &C = cs[c];
    const int ng = C.rules.ng;
    if (ctl[c].stopped || ng == 0 || it > ng) return;
    const int nb_all = bpref[ncase];
    const size_t n = (size_t)C.nx * C.ny * C.nz;
    const size_t i = (size_t)(b - bpref[c]) * 256 + threadIdx.x;
    // sub-iteration it-1's image and gradient are (u, v) when it-1 is even,
    // (u2, v2) when odd; sub-iteration it writes the other pair
    const bool odd = it & 1;
    const float *u = (odd ? u_all : u2_all) + C.voff;
    const float *vp = (odd ? v_all : v2_all) + C.voff;
    float *un = (odd ? u2_all : u_all) + C.voff;
    float *vn = (odd ? v2_all : v_all) + C.voff;
    double *prow = partials + (size_t)(it & 1) * nb_all;
    const double *prev = partials + (size_t)((it + 1) & 1) * nb_all;
    double st = 0.0;
    if (it > 0) {
        double nrm2[1];
        fold_case<1>(prev, 0, bpref[c], bpref[c + 1] - bpref[c], nrm2);
        st = ctl[c].alpha * sqrt(ctl[c].d2) / sqrt(nrm2[0]);
    }
    if (it == ng) {
        // last step (imgTVGradMin.cpp:104-116): into the case's image
        if (i < n) {
            double w = (double)u[i] - st * (double)vp[i];
            if (w < 0) w = 0;
            u_all[C.voff + i] = (float)w;
        }
        return;
    }
    // image and gradient of this sub-iteration: (u, vp) stepped when it > 0
    const float *src_u = it == 0 ? u_all + C.voff : u;
    double s = 0.0;
    if (i < n) {
        const VoxelCursor cur(i, C.nx, C.ny);
        const int x = cur.x, y = cur.y, z = cur.z;
        float g;
        if (it == 0) {
            const TvSrc<float, float, false> U{src_u, nullptr, TvStepVal{0.0, 0.0, 1.0}, C.nx, C.ny, C.nz};
            const TvAt<float, float, false> at{U, i, x, y, z};
            g = C.rules.three_d ? tv3_grad<float>(at, x, y, z) : tv2_grad<float>(at, x, y);
        } else {
            const TvSrc<float, float, true> U{u, vp, TvStepVal{st, 0.0, 1.0}, C.nx, C.ny, C.nz};
            const TvAt<float, float, true> at{U, i, x, y, z};
            g = C.rules.three_d ? tv3_grad<float>(at, x, y, z) : tv2_grad<float>(at, x, y);
            un[i] = U.stored(i);
        }
        (it == 0 ? v_all + C.voff : vn)[i] = g;
        s = (double)g * (double)g;
    }
    s = block_sum(s);
    if (threadIdx.x == 0) prow[b] = s;
}

struct BBufs {
    float *m, *mpre, *mk1, *f0, *I0;
    double *res2, *hist;
};

__global__ __launch_bounds__(256) void post_sart_batch(const BCase *__restrict__ cs,
                                                       const int *__restrict__ bpref, int ncase,
                                                       BBufs bb, const BCtl *ctl,
                                                       double *partials)
{
    const int b = blockIdx.x;
    const int c = find_seg(bpref, ncase, b);
    const BCase &C = cs[c];
    if (ctl[c].stopped) return;
    const size_t n = (size_t)C.nx * C.ny * C.nz;
    const size_t i = (size_t)(b - bpref[c]) * 256 + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < n) {
        const size_t k = C.voff + i;
        float mi = bb.m[k];
        double a = (double)mi - (double)bb.mpre[k];
        double d = (double)mi - (double)bb.mk1[k];
        v[0] = a * a;
        v[1] = d * d;
        if (C.has_I0) {
            double e = (double)mi - (double)bb.I0[k];
            v[2] = e * e;
        }
        bb.f0[k] = mi;
    }
    if (i < (size_t)C.M) {
        v[3] = bb.res2[C.roff + i];
        bb.res2[C.roff + i] = 0.0;
    }
    store_partials<4>(v, partials);
}

// One block per case: fold, stop rule, RMSE (ASD_POCS_cuda.m:136-149).
__global__ __launch_bounds__(256) void ctl_sart_batch(const BCase *__restrict__ cs,
                                                      const int *__restrict__ bpref, int ncase,
                                                      BBufs bb, BCtl *ctl, const double *partials,
                                                      int sweep_only)
{
    const int c = blockIdx.x;
    const BCase &C = cs[c];
    if (ctl[c].stopped) return;
    double tot[4];
    fold_case<4>(partials, bpref[ncase], bpref[c], bpref[c + 1] - bpref[c], tot);
    if (threadIdx.x != 0) return;
    const size_t n = (size_t)C.nx * C.ny * C.nz;
    BCtl &K = ctl[c];
    K.dd2 = tot[3];
    K.d2 = tot[1];
    if (sweep_only) return;
    const int k = K.k + 1;
    K.k = k;
    if (C.has_I0 && k - 1 < B_HIST) bb.hist[(size_t)c * B_HIST + k - 1] = sqrt(tot[2] / (double)n);
    K.dp = sqrt(tot[0]);
    if (K.dp / (double)n <= C.rules.stopRelChg) K.stopped = 1;
}

__global__ __launch_bounds__(256) void post_tv_batch(const BCase *__restrict__ cs,
                                                     const int *__restrict__ bpref, int ncase,
                                                     BBufs bb, const BCtl *ctl, double *partials)
{
    const int b = blockIdx.x;
    const int c = find_seg(bpref, ncase, b);
    const BCase &C = cs[c];
    if (ctl[c].stopped) return;
    const size_t n = (size_t)C.nx * C.ny * C.nz;
    const size_t i = (size_t)(b - bpref[c]) * 256 + threadIdx.x;
    double v[1] = {0.0};
    if (i < n) {
        const size_t k = C.voff + i;
        float mi = bb.m[k];
        double e = (double)mi - (double)bb.f0[k];
        v[0] = e * e;
        bb.mpre[k] = mi;
    }
    store_partials<1>(v, partials);
}

// One block per case: dg and the alpha / beta rules (ASD_POCS_cuda.m:166-170).
__global__ __launch_bounds__(256) void ctl_tv_batch(const BCase *__restrict__ cs,
                                                    const int *__restrict__ bpref, int ncase,
                                                    BCtl *ctl, const double *partials)
{
    const int c = blockIdx.x;
    const BCase &C = cs[c];
    if (ctl[c].stopped) return;
    double tot[1];
    fold_case<1>(partials, bpref[ncase], bpref[c], bpref[c + 1] - bpref[c], tot);
    if (threadIdx.x != 0) return;
    BCtl &K = ctl[c];
    K.dg = sqrt(tot[0]);
    if (K.dg > C.rules.r_max * K.dp && sqrt(K.dd2) > C.rules.epsilon)
        K.alpha = K.alpha * C.rules.alpha_red;
    K.beta = K.beta * C.rules.beta_red;
}

__global__ __launch_bounds__(256) void b_copy3(const float *a, float *b, float *c, float *d,
                                               size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        float x = a[i];
        b[i] = x;
        c[i] = x;
        d[i] = x;
    }
}

__global__ __launch_bounds__(256) void b_cvt_f32(const double *a, float *b, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

unsigned gridn(size_t n)
{
    size_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    return (unsigned)(b ? b : 1);
}

}  // namespace
}  // namespace asd

using namespace asd;

struct asd_batch {
    int K = 0;                 // cases
    int precision = 0;
    std::vector<asd_para> para;
    std::vector<int> nv;
    std::vector<int> order;    // cases sorted by nv (descending)
    std::vector<size_t> voff, roff, poff;
    size_t Ntot = 0, Mtot = 0, Ptot = 0;
    int max_nv = 0, max_ng = 0, lpr = 64;
    int atx_chunk = 0;                    // backprojector pairs per load batch (0: not tuned)
    double atx_us[3] = {-1.0, -1.0, -1.0};  // tuning times of chunks 1, 2, 4 (µs)
    std::vector<int> active;   // active cases per view step (prefix of `order`)
    hipStream_t stream = nullptr;
    DeviceBuf m, mpre, mk1, f0, I0, v, u2, v2, cor, res2, rt, P, ydet, zdet, ctlb, hist;
    DeviceBuf items, rpref, bpref, cases, cbpref, partials;
    std::vector<BItem> h_items;  // host copy of the per-step work lists
    std::vector<int> pos;        // position of each case in `order`
    std::vector<int> h_rpref_tot, h_bpref_tot;  // per step: total rays / blocks
    int case_blocks = 0;
    std::vector<char> have_P, have_I0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool use_graph = true;
    std::vector<BCtl> h;
};

namespace {

int b_enqueue_iteration(asd_batch *B, int sweep_only, int iterate)
{
    const int K = B->K;
    // sweep: view steps
    for (int j = 0; j < B->max_nv; j++) {
        const int A = B->active[j];
        if (A == 0) break;
        const BItem *items = B->items.as<BItem>() + (size_t)j * K;
        const int *rp = B->rpref.as<int>() + (size_t)j * (K + 1);
        const int *bp = B->bpref.as<int>() + (size_t)j * (K + 1);
        const int blocks = B->h_bpref_tot[j];
        const unsigned ab = (unsigned)B->h_rpref_tot[j];  // forward-projector blocks
        if (B->precision == 1) {
            switch (B->lpr) {
            case 64: ax_batch_kernel<64, true, SinoU16><<<ab, 256, 0, B->stream>>>(items, rp, A); break;
            case 16: ax_batch_kernel<16, true, SinoU16><<<ab, 256, 0, B->stream>>>(items, rp, A); break;
            case 4: ax_batch_kernel<4, true, SinoU16><<<ab, 256, 0, B->stream>>>(items, rp, A); break;
            default: ax_batch_kernel<1, true, SinoU16><<<ab, 256, 0, B->stream>>>(items, rp, A); break;
            }
            ASD_LAUNCH_CHECK("ax_batch_kernel");
            launch_atx_batch<true>(B->atx_chunk, blocks, B->stream, items, bp, A, 0);
        } else {
            switch (B->lpr) {
            case 64: ax_batch_kernel<64, false, SinoF32><<<ab, 256, 0, B->stream>>>(items, rp, A); break;
            case 16: ax_batch_kernel<16, false, SinoF32><<<ab, 256, 0, B->stream>>>(items, rp, A); break;
            case 4: ax_batch_kernel<4, false, SinoF32><<<ab, 256, 0, B->stream>>>(items, rp, A); break;
            default: ax_batch_kernel<1, false, SinoF32><<<ab, 256, 0, B->stream>>>(items, rp, A); break;
            }
            ASD_LAUNCH_CHECK("ax_batch_kernel");
            launch_atx_batch<false>(B->atx_chunk, blocks, B->stream, items, bp, A, 0);
        }
        ASD_LAUNCH_CHECK("atx_batch_kernel");
    }
    BBufs bb{B->m.as<float>(), B->mpre.as<float>(), B->mk1.as<float>(), B->f0.as<float>(),
             B->I0.as<float>(), B->res2.as<double>(), B->hist.as<double>()};
    const BCase *cs = B->cases.as<BCase>();
    const int *cbp = B->cbpref.as<int>();
    BCtl *ctl = B->ctlb.as<BCtl>();
    double *part = B->partials.as<double>();
    post_sart_batch<<<B->case_blocks, 256, 0, B->stream>>>(cs, cbp, K, bb, ctl, part);
    ASD_LAUNCH_CHECK("post_sart_batch");
    ctl_sart_batch<<<K, 256, 0, B->stream>>>(cs, cbp, K, bb, ctl, part, sweep_only);
    ASD_LAUNCH_CHECK("ctl_sart_batch");
    if (!iterate) return ASD_OK;
    for (int it = 0; it <= B->max_ng; it++) {
        tv_sub_batch<<<B->case_blocks, 256, 0, B->stream>>>(cs, cbp, K, B->m.as<float>(),
                                                            B->v.as<float>(), B->u2.as<float>(),
                                                            B->v2.as<float>(), ctl, part, it);
        ASD_LAUNCH_CHECK("tv_sub_batch");
    }
    post_tv_batch<<<B->case_blocks, 256, 0, B->stream>>>(cs, cbp, K, bb, ctl, part);
    ASD_LAUNCH_CHECK("post_tv_batch");
    ctl_tv_batch<<<K, 256, 0, B->stream>>>(cs, cbp, K, ctl, part);
    ASD_LAUNCH_CHECK("ctl_tv_batch");
    return ASD_OK;
}

// Picks the backprojector's pairs per load batch: each choice runs the first
// view step (every case active) dry, timed by the dispatch itself, best of
// three after a warm-up; a choice must beat the default (2) by more than 3%,
// so near-ties do not flip between runs.  All choices give the same bits.
int b_autotune(asd_batch *B)
{
    const int A = B->active[0];
    const BItem *items = B->items.as<BItem>();
    const int *bp = B->bpref.as<int>();
    const unsigned blocks = (unsigned)B->h_bpref_tot[0];
    ASD_HIP(hipMemsetAsync(B->cor.p, 0, B->Mtot * 4, B->stream));
    hipEvent_t e0, e1;
    ASD_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        set_error(ASD_ERR_HIP, "hipEventCreate failed");
        return ASD_ERR_HIP;
    }
    const int chunks[3] = {2, 1, 4};
    int best = 2;
    double best_us = 1e30;
    int rc = ASD_OK;
    for (int chunk : chunks) {
        float ms_min = 1e30f;
        for (int rep = 0; rep < 4 && rc == ASD_OK; rep++) {
            LaunchTiming lt{e0, e1, 0};
            t_launch_timing = &lt;
            if (B->precision == 1)
                launch_atx_batch<true>(chunk, blocks, B->stream, items, bp, A, 1);
            else
                launch_atx_batch<false>(chunk, blocks, B->stream, items, bp, A, 1);
            t_launch_timing = nullptr;
            if (hipGetLastError() != hipSuccess) {
                set_error(ASD_ERR_HIP, "atx_batch_kernel launch failed");
                rc = ASD_ERR_HIP;
                break;
            }
            (void)hipEventSynchronize(e1);
            float ms = 0;
            if (rep > 0 && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms < ms_min)
                ms_min = ms;
        }
        if (rc) break;
        B->atx_us[chunk == 4 ? 2 : chunk - 1] = ms_min * 1e3;
        if (ms_min * 1e3 < best_us * 0.97) {
            best_us = ms_min * 1e3;
            best = chunk;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    B->atx_chunk = best;
    return ASD_OK;
}

int b_sync(asd_batch *B)
{
    ASD_HIP(hipMemcpyAsync(B->h.data(), B->ctlb.p, sizeof(BCtl) * B->K, hipMemcpyDeviceToHost,
                           B->stream));
    ASD_HIP(hipStreamSynchronize(B->stream));
    return ASD_OK;
}

// As the single plan's drop_graph: the stream is drained before the
// executable graph is destroyed.
void b_drop_graph(asd_batch *B)
{
    if (B->exec && B->stream) (void)hipStreamSynchronize(B->stream);
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    B->exec = nullptr;
    B->graph = nullptr;
}

}  // namespace

extern "C" {

void asd_batch_destroy(asd_batch *B)
{
    if (!B) return;
    if (B->stream) (void)hipStreamSynchronize(B->stream);
    b_drop_graph(B);
    if (B->stream) (void)hipStreamDestroy(B->stream);
    delete B;
}

int asd_batch_create(int ncases, const asd_para *paras, const double *const *phi_vecs,
                     const int *nvs, asd_batch **out)
{
    clear_error();
    ASD_ARG(out && paras && phi_vecs && nvs, "NULL argument");
    ASD_ARG(ncases > 0, "ncases must be positive");
    *out = nullptr;
    for (int c = 0; c < ncases; c++) {
        const asd_para &p = paras[c];
        ASD_ARG(p.nx > 0 && p.ny > 0 && p.nz > 0 && p.na > 0 && p.nb > 0,
                "every case needs positive nx, ny, nz, na, nb");
        ASD_ARG(index_fits(p.nx, p.ny, p.nz, p.na, p.nb), "case " ASD_INDEX_LIMITS);
        ASD_ARG(vol_buffer_ok((size_t)p.nx * p.ny * p.nz, p.nx),
                "batched cases need fewer than 2^30 voxels and nx below 2^22");
        ASD_ARG(nvs[c] > 0 && phi_vecs[c], "every case needs views");
        ASD_ARG(p.SO > 0 && p.OD > 0 && p.vxlsize > 0 && p.detsize > 0, "bad case geometry");
        ASD_ARG(p.precision == paras[0].precision, "all cases must share para.precision");
        ASD_ARG(p.precision == 0 || p.precision == 1, "batched para.precision must be 0 or 1");
        ASD_ARG(p.numOfgradients >= 0, "para.numOfgradients must be >= 0");
    }
    asd_batch *B = new asd_batch();
    auto fail = [&](int code) { asd_batch_destroy(B); return code; };
    const int K = ncases;
    B->K = K;
    B->precision = paras[0].precision;
    B->para.assign(paras, paras + K);
    B->nv.assign(nvs, nvs + K);
    B->voff.resize(K);
    B->roff.resize(K);
    B->poff.resize(K);
    long maxrays = 0;
    int maxu = 0;
    for (int c = 0; c < K; c++) {
        const asd_para &p = paras[c];
        B->voff[c] = B->Ntot;
        B->roff[c] = B->Mtot;
        B->poff[c] = B->Ptot;
        B->Ntot += (size_t)p.nx * p.ny * p.nz;
        B->Mtot += (size_t)p.na * p.nb;
        B->Ptot += (size_t)p.na * p.nb * nvs[c];
        B->max_nv = std::max(B->max_nv, nvs[c]);
        B->max_ng = std::max(B->max_ng, p.numOfgradients);
        maxrays = std::max(maxrays, (long)p.na * p.nb);
        maxu = std::max(maxu, std::max(p.nx, p.ny));
    }
    // lanes per ray from the largest case, restricted to the compiled set
    int l = ax_lanes_per_ray((int)std::min<long>(maxrays, 1L << 30), 1, maxu);
    B->lpr = l >= 64 ? 64 : l >= 16 ? 16 : l >= 4 ? 4 : 1;
    if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess) {
        B->stream = nullptr;
        set_error(ASD_ERR_HIP, "hipStreamCreateWithFlags failed");
        return fail(ASD_ERR_HIP);
    }
    B->order.resize(K);
    std::iota(B->order.begin(), B->order.end(), 0);
    std::stable_sort(B->order.begin(), B->order.end(),
                     [&](int a, int b) { return nvs[a] > nvs[b]; });
    // per-case device buffers
    const size_t N = B->Ntot, M = B->Mtot;
    std::vector<size_t> yoff(K), zoff(K);
    size_t ytot = 0, ztot = 0;
    for (int c = 0; c < K; c++) {
        yoff[c] = ytot;
        zoff[c] = ztot;
        ytot += (size_t)paras[c].na * nvs[c];
        ztot += (size_t)paras[c].nb * nvs[c];
    }
    int rc;
    if ((rc = B->m.ensure(N * 4)) || (rc = B->mpre.ensure(N * 4)) || (rc = B->mk1.ensure(N * 4)) ||
        (rc = B->f0.ensure(N * 4)) || (rc = B->I0.ensure(N * 4)) || (rc = B->v.ensure(N * 4)) ||
        (rc = B->u2.ensure(N * 4)) || (rc = B->v2.ensure(N * 4)) ||
        (rc = B->cor.ensure(M * 4)) || (rc = B->res2.ensure(M * 8)) ||
        (rc = B->rt.ensure(B->Ptot * sizeof(RayRec))) ||
        (rc = B->P.ensure(B->Ptot * sino_bytes(B->precision))) ||
        (rc = B->ydet.ensure(ytot * 4)) || (rc = B->zdet.ensure(ztot * 4)) ||
        (rc = B->ctlb.ensure(sizeof(BCtl) * K)) || (rc = B->hist.ensure((size_t)K * B_HIST * 8)))
        return fail(rc);
    // geometry tables and per-step work lists
    std::vector<float> yall(ytot), zall(ztot);
    std::vector<std::vector<ViewGeom>> geoms(K);
    for (int c = 0; c < K; c++) {
        const asd_para &p = paras[c];
        geoms[c].resize(nvs[c]);
        for (int j = 0; j < nvs[c]; j++) {
            ViewGeom &g = geoms[c][j];
            double theta = phi_vecs[c][j] / 180 * M_PI;
            float phi = (float)(-1 * theta);
            float sc = (float)p.vxlsize, det = (float)p.detsize / sc;
            g.cos_phi = (float)std::cos((double)phi);
            g.sin_phi = (float)std::sin((double)phi);
            g.SO = (float)p.SO / sc;
            g.OD = (float)p.OD / sc;
            g.scale = sc;
            g.dz = 1.0f;
            g.dy_det = g.dz_det = det;
            g.nx = p.nx; g.ny = p.ny; g.nz = p.nz; g.na = p.na; g.nb = p.nb;
            geom_finish(g);
            for (int i = 0; i < p.na; i++)
                yall[yoff[c] + (size_t)j * p.na + i] =
                    (float)(((double)(float)i - (double)(float)p.na / 2.0 + 0.5) * (double)det);
            for (int i = 0; i < p.nb; i++)
                zall[zoff[c] + (size_t)j * p.nb + i] =
                    (float)(((double)(float)i - (double)(float)p.nb / 2.0 + 0.5) * (double)det);
            g.tight = view_tight(g, &yall[yoff[c] + (size_t)j * p.na]);
            g.general = view_general(g, &yall[yoff[c] + (size_t)j * p.na],
                                     &zall[zoff[c] + (size_t)j * p.nb]);
        }
    }
    std::vector<BItem> items((size_t)B->max_nv * K);
    std::vector<int> rp((size_t)B->max_nv * (K + 1), 0), bp((size_t)B->max_nv * (K + 1), 0);
    B->active.assign(B->max_nv, 0);
    B->h_rpref_tot.assign(B->max_nv, 0);
    B->h_bpref_tot.assign(B->max_nv, 0);
    BCtl *dctl = B->ctlb.as<BCtl>();
    const size_t Pe = sino_bytes(B->precision);
    const int rpb = 256 / B->lpr;  // rays per forward-projector block
    for (int j = 0; j < B->max_nv; j++) {
        int A = 0, rays = 0, blocks = 0;
        for (int s = 0; s < K; s++) {
            int c = B->order[s];
            if (nvs[c] <= j) break;
            const asd_para &p = paras[c];
            BItem &it = items[(size_t)j * K + A];
            it.g = geoms[c][j];
            it.m = B->m.as<float>() + B->voff[c];
            it.P = (const char *)B->P.p + (B->poff[c] + (size_t)j * p.na * p.nb) * Pe;
            it.cor = B->cor.as<float>() + B->roff[c];
            it.res2 = B->res2.as<double>() + B->roff[c];
            it.rt = B->rt.as<RayRec>() + B->poff[c] + (size_t)j * p.na * p.nb;
            it.ydet = B->ydet.as<float>() + yoff[c] + (size_t)j * p.na;
            it.zdet = B->zdet.as<float>() + zoff[c] + (size_t)j * p.nb;
            it.ctl = dctl + c;
            it.lo = 0.0;
            it.step = 0.0;
            rp[(size_t)j * (K + 1) + A] = rays;
            bp[(size_t)j * (K + 1) + A] = blocks;
            rays += (p.na * p.nb + rpb - 1) / rpb;  // whole blocks per case
            blocks += (int)(((size_t)p.nx * p.ny * p.nz + 255) / 256);
            A++;
        }
        rp[(size_t)j * (K + 1) + A] = rays;
        bp[(size_t)j * (K + 1) + A] = blocks;
        B->active[j] = A;
        B->h_rpref_tot[j] = rays;
        B->h_bpref_tot[j] = blocks;
    }
    std::vector<BCase> cs(K);
    std::vector<int> cbp(K + 1, 0);
    for (int c = 0; c < K; c++) {
        const asd_para &p = paras[c];
        cs[c].nx = p.nx; cs[c].ny = p.ny; cs[c].nz = p.nz;
        cs[c].M = p.na * p.nb;
        cs[c].voff = B->voff[c];
        cs[c].roff = B->roff[c];
        cs[c].rules = BRules{p.alpha_red, p.beta_red, p.r_max, p.epsilon, p.stopRelChg,
                             p.numOfgradients, p.nz >= 2 ? 1 : 0};
        cs[c].has_I0 = 0;
        size_t need = std::max((size_t)p.nx * p.ny * p.nz, (size_t)p.na * p.nb);
        cbp[c + 1] = cbp[c] + (int)((need + 255) / 256);
    }
    B->case_blocks = cbp[K];
    B->h_items = items;
    B->pos.assign(K, 0);
    for (int s2 = 0; s2 < K; s2++) B->pos[B->order[s2]] = s2;
    if ((rc = B->items.ensure(items.size() * sizeof(BItem))) ||
        (rc = B->rpref.ensure(rp.size() * 4)) || (rc = B->bpref.ensure(bp.size() * 4)) ||
        (rc = B->cases.ensure(cs.size() * sizeof(BCase))) || (rc = B->cbpref.ensure(cbp.size() * 4)) ||
        (rc = B->partials.ensure((size_t)4 * B->case_blocks * 8)))
        return fail(rc);
    if (hipMemcpy(B->ydet.p, yall.data(), ytot * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(B->zdet.p, zall.data(), ztot * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(B->items.p, items.data(), items.size() * sizeof(BItem), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(B->rpref.p, rp.data(), rp.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(B->bpref.p, bp.data(), bp.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(B->cases.p, cs.data(), cs.size() * sizeof(BCase), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(B->cbpref.p, cbp.data(), cbp.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(B->m.p, 0, N * 4) != hipSuccess || hipMemset(B->res2.p, 0, M * 8) != hipSuccess ||
        hipMemset(B->I0.p, 0, N * 4) != hipSuccess || hipMemset(B->P.p, 0, B->P.bytes) != hipSuccess) {
        set_error(ASD_ERR_HIP, "batch initialisation failed");
        return fail(ASD_ERR_HIP);
    }
    B->have_P.assign(K, 0);
    B->have_I0.assign(K, 0);
    B->h.assign(K, BCtl{});
    for (int c = 0; c < K; c++) {
        B->h[c].alpha = paras[c].alpha;
        B->h[c].beta = paras[c].beta;
    }
    if (hipMemcpy(B->ctlb.p, B->h.data(), sizeof(BCtl) * K, hipMemcpyHostToDevice) != hipSuccess) {
        set_error(ASD_ERR_HIP, "batch initialisation failed");
        return fail(ASD_ERR_HIP);
    }
    // the ray-slope tables of every case and view, once
    for (int c = 0; c < K; c++)
        for (int j = 0; j < nvs[c]; j++) {
            const asd_para &p = paras[c];
            if ((rc = launch_ray_table(geoms[c][j], B->ydet.as<float>() + yoff[c] + (size_t)j * p.na,
                                       B->zdet.as<float>() + zoff[c] + (size_t)j * p.nb,
                                       B->rt.as<RayRec>() + B->poff[c] + (size_t)j * p.na * p.nb,
                                       B->stream)))
                return fail(rc);
        }
    // start images: zeros
    b_copy3<<<gridn(N), 256, 0, B->stream>>>(B->m.as<float>(), B->mpre.as<float>(),
                                             B->mk1.as<float>(), B->f0.as<float>(), N);
    if (hipStreamSynchronize(B->stream) != hipSuccess) {
        set_error(ASD_ERR_HIP, "batch initialisation failed");
        return fail(ASD_ERR_HIP);
    }
    *out = B;
    return ASD_OK;
}

int asd_batch_set_case(asd_batch *B, int c, const double *P, const double *m0, const double *I0)
{
    clear_error();
    ASD_ARG(B && P, "NULL argument");
    ASD_ARG(c >= 0 && c < B->K, "case index out of range");
    const asd_para &p = B->para[c];
    const size_t np = (size_t)p.na * p.nb * B->nv[c], n = (size_t)p.nx * p.ny * p.nz;
    int rc;
    DeviceBuf tmp;
    if ((rc = tmp.ensure(std::max(np, n) * 8))) return rc;
    if (B->precision == 1) {
        // 16-bit fixed point per view (see sino_codec.hpp); the view scales go
        // into this case's work-list entries
        const size_t Mc = (size_t)p.na * p.nb;
        std::vector<uint16_t> q(np);
        for (int j = 0; j < B->nv[c]; j++) {
            const double *v = P + (size_t)j * Mc;
            double lo = v[0], hi = v[0];
            for (size_t i = 1; i < Mc; i++) {
                lo = v[i] < lo ? v[i] : lo;
                hi = v[i] > hi ? v[i] : hi;
            }
            const double step = hi > lo ? (hi - lo) / 65535.0 : 0.0;
            for (size_t i = 0; i < Mc; i++) {
                double t = step > 0 ? (v[i] - lo) / step : 0.0;
                long r = std::lround(t);
                q[(size_t)j * Mc + i] = (uint16_t)(r < 0 ? 0 : r > 65535 ? 65535 : r);
            }
            BItem &it = B->h_items[(size_t)j * B->K + B->pos[c]];
            it.lo = lo;
            it.step = step;
            ASD_HIP(hipMemcpyAsync(B->items.as<BItem>() + (size_t)j * B->K + B->pos[c], &it,
                                   sizeof(BItem), hipMemcpyHostToDevice, B->stream));
        }
        ASD_HIP(hipMemcpyAsync(B->P.as<uint16_t>() + B->poff[c], q.data(), np * 2,
                               hipMemcpyHostToDevice, B->stream));
        ASD_HIP(hipStreamSynchronize(B->stream));
    } else {
        ASD_HIP(hipMemcpyAsync(tmp.p, P, np * 8, hipMemcpyHostToDevice, B->stream));
        b_cvt_f32<<<gridn(np), 256, 0, B->stream>>>(tmp.as<double>(), B->P.as<float>() + B->poff[c], np);
        ASD_LAUNCH_CHECK("b_cvt");
        ASD_HIP(hipStreamSynchronize(B->stream));
    }
    float *mc = B->m.as<float>() + B->voff[c];
    if (m0) {
        ASD_HIP(hipMemcpyAsync(tmp.p, m0, n * 8, hipMemcpyHostToDevice, B->stream));
        b_cvt_f32<<<gridn(n), 256, 0, B->stream>>>(tmp.as<double>(), mc, n);
        ASD_LAUNCH_CHECK("b_cvt");
    } else {
        ASD_HIP(hipMemsetAsync(mc, 0, n * 4, B->stream));
    }
    b_copy3<<<gridn(n), 256, 0, B->stream>>>(mc, B->mpre.as<float>() + B->voff[c],
                                             B->mk1.as<float>() + B->voff[c],
                                             B->f0.as<float>() + B->voff[c], n);
    ASD_LAUNCH_CHECK("b_copy3");
    int has = 0;
    if (I0) {
        ASD_HIP(hipStreamSynchronize(B->stream));
        ASD_HIP(hipMemcpyAsync(tmp.p, I0, n * 8, hipMemcpyHostToDevice, B->stream));
        b_cvt_f32<<<gridn(n), 256, 0, B->stream>>>(tmp.as<double>(), B->I0.as<float>() + B->voff[c], n);
        ASD_LAUNCH_CHECK("b_cvt");
        has = 1;
    }
    ASD_HIP(hipMemcpyAsync((char *)B->cases.p + sizeof(BCase) * c + offsetof(BCase, has_I0), &has,
                           sizeof(int), hipMemcpyHostToDevice, B->stream));
    B->h[c] = BCtl{};
    B->h[c].alpha = p.alpha;
    B->h[c].beta = p.beta;
    ASD_HIP(hipMemcpyAsync(B->ctlb.as<BCtl>() + c, &B->h[c], sizeof(BCtl), hipMemcpyHostToDevice,
                           B->stream));
    ASD_HIP(hipStreamSynchronize(B->stream));
    B->have_P[c] = 1;
    B->have_I0[c] = (char)has;
    return ASD_OK;
}

int asd_batch_iterate(asd_batch *B, int n_iter)
{
    clear_error();
    ASD_ARG(B, "NULL batch");
    ASD_ARG(n_iter >= 0, "n_iter must be >= 0");
    for (int c = 0; c < B->K; c++)
        if (!B->have_P[c]) {
            set_error(ASD_ERR_STATE, "case " + std::to_string(c) + " has no projections");
            return ASD_ERR_STATE;
        }
    int rc;
    if (!B->atx_chunk && (rc = b_autotune(B))) return rc;
    if (B->use_graph) {
        if (!B->exec) {
            ASD_HIP(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
            rc = b_enqueue_iteration(B, 0, 1);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(B->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) {
                set_error(ASD_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
                return ASD_ERR_HIP;
            }
            hipError_t ei = hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0);
            if (ei != hipSuccess) {
                (void)hipGraphDestroy(g);
                B->exec = nullptr;
                set_error(ASD_ERR_HIP, std::string("hipGraphInstantiate: ") +
                                           hipGetErrorString(ei));
                return ASD_ERR_HIP;
            }
            B->graph = g;
        }
        for (int i = 0; i < n_iter; i++) ASD_HIP(hipGraphLaunch(B->exec, B->stream));
    } else {
        for (int i = 0; i < n_iter; i++)
            if ((rc = b_enqueue_iteration(B, 0, 1))) return rc;
    }
    return b_sync(B);
}

int asd_batch_set_graph(asd_batch *B, int enable)
{
    clear_error();
    ASD_ARG(B, "NULL batch");
    B->use_graph = enable != 0;
    if (!B->use_graph) b_drop_graph(B);
    return ASD_OK;
}

int asd_batch_set_atx_chunk(asd_batch *B, int chunk)
{
    clear_error();
    ASD_ARG(B, "NULL batch");
    ASD_ARG(chunk == 0 || chunk == 1 || chunk == 2 || chunk == 4, "chunk must be 0, 1, 2 or 4");
    b_drop_graph(B);
    B->atx_chunk = chunk;
    return ASD_OK;
}

int asd_batch_tuning(asd_batch *B, int *chunk, double *us3)
{
    clear_error();
    ASD_ARG(B, "NULL batch");
    if (chunk) *chunk = B->atx_chunk;
    if (us3)
        for (int k = 0; k < 3; k++) us3[k] = B->atx_us[k];
    return ASD_OK;
}

int asd_batch_sync(asd_batch *B)
{
    clear_error();
    ASD_ARG(B, "NULL batch");
    ASD_HIP(hipStreamSynchronize(B->stream));
    return ASD_OK;
}

int asd_batch_get_image(asd_batch *B, int c, double *m_res)
{
    clear_error();
    ASD_ARG(B && m_res, "NULL argument");
    ASD_ARG(c >= 0 && c < B->K, "case index out of range");
    const asd_para &p = B->para[c];
    size_t n = (size_t)p.nx * p.ny * p.nz;
    std::vector<float> h(n);
    ASD_HIP(hipMemcpyAsync(h.data(), B->f0.as<float>() + B->voff[c], n * 4, hipMemcpyDeviceToHost,
                           B->stream));
    ASD_HIP(hipStreamSynchronize(B->stream));
    for (size_t i = 0; i < n; i++) m_res[i] = h[i];
    return ASD_OK;
}

int asd_batch_get_out(asd_batch *B, int c, asd_out *o, double *rmse)
{
    clear_error();
    ASD_ARG(B && o, "NULL argument");
    ASD_ARG(c >= 0 && c < B->K, "case index out of range");
    const BCtl &h = B->h[c];
    o->IterationNum = h.k;
    o->dataTolerance = h.dd2;
    o->alpha = h.alpha;
    o->beta = h.beta;
    if (rmse && B->have_I0[c] && h.k > 0) {
        int n = h.k < B_HIST ? h.k : B_HIST;
        ASD_HIP(hipMemcpy(rmse, B->hist.as<double>() + (size_t)c * B_HIST, n * 8,
                          hipMemcpyDeviceToHost));
    }
    return ASD_OK;
}

}  // extern "C"
