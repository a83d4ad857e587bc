// Device-resident ASD-POCS driver: the loop of utilities/ASD_POCS_cuda.m:103-171
// and the sweep of SART_cuda_inline.m:56-87 with the image, the sinogram and
// every per-view quantity kept in HBM.
//
// One outer iteration is a fixed sequence of kernels with no host decision
// in it, so it can be captured once into a hipGraph and replayed:
//   per view   ax_sart    ray sums -> cor = (P - vxl*y) / (vxl*sr | 1), the
//                         per-ray squared residual and the view's slope table
//                         (ASD_POCS_cuda.m:117-123)
//              atx_sart   voxel sums -> m = max(0, m + beta*(vxl*x)/(vxl*sc | 1))
//                         (ASD_POCS_cuda.m:125-131)
//   post_sart  dp = ||m - m_pre||, d = ||m - m_k1||, RMSE, dd2; f0 = m; the
//              stop rule (m:136-149)
//   TV         ng x (gradient + FP64 norm, normalised step) (m:155-164)
//   post_tv    dg = ||m - f0||; m_pre = m; the alpha and beta rules (m:166-170)
// alpha, beta, k and the stop flag live in device memory (DevCtl); once the
// stop rule fires every later kernel returns at entry.
#include <hip/hip_fp16.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "plan.hpp"
#include "projector_kernels.hpp"
#include "reduce_device.hpp"
#include "sino_codec.hpp"
#include "tv.hpp"

namespace asd {

namespace {


template <class Codec>
struct EpiSartAx {
    Codec P;              // this view's measured projection
    float *cor;           // correction fed to the backprojector
    double *res2;         // per-ray accumulated squared residual of the sweep
    const int *stopped;
    float scale;
    struct Pre {
        double p, r2;
    };
    __device__ __forceinline__ bool skip() const { return *stopped != 0; }
    // issued before the ray march so the loads overlap it
    __device__ __forceinline__ Pre load(long id) const { return Pre{P.at(id), res2[id]}; }
    __device__ __forceinline__ void operator()(long id, float acc, float accl, const Pre &q) const
    {
#pragma clang fp contract(off)
        float proj = acc * scale, sumr = accl * scale;
        double r = (sumr == 0.0f) ? 1.0 : (double)sumr;
        double c0 = q.p - (double)proj;
        res2[id] = q.r2 + c0 * c0;
        cor[id] = (float)(c0 / r);
    }
};

struct EpiSartAtx {
    float *m;
    const double *relax;  // beta of the ASD loop or lambda of a stand-alone sweep
    const int *stopped;
    float scale;
    struct Pre {
        float m;
    };
    __device__ __forceinline__ bool skip() const { return *stopped != 0; }
    __device__ __forceinline__ Pre load(long v) const { return Pre{m[v]}; }
    __device__ __forceinline__ void operator()(long v, float acc, float accl, const Pre &q) const
    {
#pragma clang fp contract(off)
        float bp = acc * scale, sc = accl * scale;
        double c = (sc == 0.0f) ? 1.0 : (double)sc;
        double nm = (double)q.m + (*relax) * ((double)bp / c);
        m[v] = nm < 0 ? 0.0f : (float)nm;
    }
};

// The update of the reference-arithmetic plan: m in FP64 as MATLAB keeps it,
// m = m + beta * (bProj ./ sumC), then m(m < 0) = 0 (ASD_POCS_cuda.m:126-131),
// and its FP32 shadow ms = single(m), which the next view's forward
// projector reads (m:117).
struct EpiSartAtx64 {
    double *m;
    float *ms;
    const double *relax;
    const int *stopped;
    float scale;
    struct Pre {
        double m;
    };
    __device__ __forceinline__ bool skip() const { return *stopped != 0; }
    __device__ __forceinline__ Pre load(long v) const { return Pre{m[v]}; }
    __device__ __forceinline__ void operator()(long v, float acc, float accl, const Pre &q) const
    {
#pragma clang fp contract(off)
        float bp = acc * scale, sc = accl * scale;
        double c = (sc == 0.0f) ? 1.0 : (double)sc;
        double nm = q.m + (*relax) * ((double)bp / c);
        nm = nm < 0 ? 0.0 : nm;
        m[v] = nm;
        ms[v] = (float)nm;
    }
};

// After a sweep: dp^2, d^2, rmse^2 over the image (and f0 = m) and dd2 over the
// rays (and the residual buffer re-zeroed), as block partials.  T: the image
// type (FP32, or FP64 in the reference-arithmetic plan).
template <class T>
__global__ __launch_bounds__(256) void post_sart_kernel(
    const T *__restrict__ m, const T *__restrict__ mpre, const T *__restrict__ mk1,
    const T *__restrict__ I0, T *__restrict__ f0, double *__restrict__ res2, size_t N,
    size_t M, double *partials, const DevCtl *ctl)
{
    if (ctl->stopped) return;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        T mi = m[i];
        double a = (double)mi - (double)mpre[i];
        double b = (double)mi - (double)mk1[i];
        v[0] += a * a;
        v[1] += b * b;
        if (I0) {
            double c = (double)mi - (double)I0[i];
            v[2] += c * c;
        }
        f0[i] = mi;
    }
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < M; j += stride) {
        v[3] += res2[j];
        res2[j] = 0.0;
    }
    store_partials<4>(v, partials);
}

// Folds post_sart's partials; applies the stop rule and records the RMSE
// (ASD_POCS_cuda.m:136-149).  One block.  serial: dp^2, d^2 and rmse^2 are
// already in scal (summed in index order) and replace the folded ones.
__global__ __launch_bounds__(RED_THREADS) void ctl_after_sart(const double *partials, int nblocks,
                                                              size_t N, int has_I0, double *scal,
                                                              DevCtl *ctl, double *hist,
                                                              CtlRules rules, int sweep_only,
                                                              int serial)
{
    if (ctl->stopped) return;
    double tot[4];
    fold_rows<4>(partials, nblocks, tot);
    if (threadIdx.x != 0) return;
    if (serial) {
        tot[0] = scal[TV_SLOT_DP2];
        tot[1] = scal[TV_SLOT_D2];
        tot[2] = scal[TV_SLOT_RMSE2];
    }
    scal[TV_SLOT_DP2] = tot[0];
    scal[TV_SLOT_D2] = tot[1];
    scal[TV_SLOT_RMSE2] = tot[2];
    scal[TV_SLOT_DD2] = tot[3];
    ctl->dd2 = tot[3];
    if (sweep_only) return;
    const int k = ctl->k + 1;
    ctl->k = k;
    if (has_I0 && k - 1 < HIST_CAP) hist[k - 1] = sqrt(tot[2] / (double)N);
    const double dp = sqrt(tot[0]);
    ctl->dp = dp;
    if (dp / (double)N <= rules.stopRelChg) ctl->stopped = 1;
}

// After TV: dg^2 = ||m - f0||^2 partials and m_pre = m for the next
// iteration; ms (when not NULL) receives single(m) for the next sweep.
template <class T>
__global__ __launch_bounds__(256) void post_tv_kernel(const T *__restrict__ m,
                                                      const T *__restrict__ f0,
                                                      T *__restrict__ mpre, float *__restrict__ ms,
                                                      size_t N, double *partials,
                                                      const DevCtl *ctl)
{
    if (ctl->stopped) return;
    double v[1] = {0.0};
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        T mi = m[i];
        double e = (double)mi - (double)f0[i];
        v[0] += e * e;
        mpre[i] = mi;
        if (ms) ms[i] = (float)mi;
    }
    store_partials<1>(v, partials);
}

// Folds dg^2 and applies the alpha and beta rules (ASD_POCS_cuda.m:166-170).
// serial: dg^2 is already in scal (summed in index order).
__global__ __launch_bounds__(RED_THREADS) void ctl_after_tv(const double *partials, int nblocks,
                                                            double *scal, DevCtl *ctl,
                                                            CtlRules rules, int serial)
{
    if (ctl->stopped) return;
    double tot[1];
    fold_rows<1>(partials, nblocks, tot);
    if (threadIdx.x != 0) return;
    if (serial) tot[0] = scal[TV_SLOT_DG2];
    scal[TV_SLOT_DG2] = tot[0];
    const double dg = sqrt(tot[0]);
    ctl->dg = dg;
    if (dg > rules.r_max * ctl->dp && sqrt(ctl->dd2) > rules.epsilon)
        ctl->alpha = ctl->alpha * rules.alpha_red;
    ctl->beta = ctl->beta * rules.beta_red;
}

__global__ __launch_bounds__(256) void f64_to_f32(const double *__restrict__ a, float *b,
                                                  size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}


template <class T>
__global__ __launch_bounds__(256) void copy3_kernel(const T *__restrict__ a, T *b, T *c, T *d,
                                                    size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        T x = a[i];
        b[i] = x;
        c[i] = x;
        d[i] = x;
    }
}

unsigned grid_for(size_t n)
{
    size_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    return (unsigned)(b ? b : 1);
}

}  // namespace
}  // namespace asd

using namespace asd;


namespace {

void view_geom_from_para(const asd_para &p, double phi_deg, ViewGeom &g, std::vector<float> &ydet,
                         std::vector<float> &zdet)
{
    // Same float casts and formulas as the MEX gateways
    // (src/AxAtx/Ax_main_test.cpp:138-161) after ASD_POCS_cuda.m:111,116.
    double theta = phi_deg / 180 * M_PI;
    float phi = (float)(-1 * theta);
    float so = (float)p.SO, od = (float)p.OD, sc = (float)p.vxlsize, det = (float)p.detsize;
    g.cos_phi = (float)std::cos((double)phi);
    g.sin_phi = (float)std::sin((double)phi);
    g.SO = so / sc;
    g.OD = od / sc;
    det = det / sc;
    g.scale = sc;
    g.dz = 1.0f;
    g.dy_det = det;
    g.dz_det = det;
    g.nx = p.nx; g.ny = p.ny; g.nz = p.nz; g.na = p.na; g.nb = p.nb;
    geom_finish(g);
    ydet.resize(p.na);
    zdet.resize(p.nb);
    for (int i = 0; i < p.na; i++)
        ydet[i] = (float)(((double)(float)i - (double)(float)p.na / 2.0 + 0.5) * (double)det);
    for (int j = 0; j < p.nb; j++)
        zdet[j] = (float)(((double)(float)j - (double)(float)p.nb / 2.0 + 0.5) * (double)det);
    g.tight = view_tight(g, ydet.data());
    g.general = view_general(g, ydet.data(), zdet.data());
}

int check_para(const asd_para *p, int nv)
{
    ASD_ARG(p, "NULL para");
    ASD_ARG(p->nx > 0 && p->ny > 0 && p->nz > 0, "para.nx, para.ny, para.nz must be positive");
    ASD_ARG(p->na > 0 && p->nb > 0, "para.na, para.nb must be positive");
    ASD_ARG(index_fits(p->nx, p->ny, p->nz, p->na, p->nb), ASD_INDEX_LIMITS);
    ASD_ARG(nv > 0, "phi_vec is empty");
    ASD_ARG(p->vxlsize > 0 && p->detsize > 0, "para.vxlsize and para.detsize must be positive");
    ASD_ARG(p->SO > 0 && p->OD > 0, "para.SO and para.OD must be positive");
    ASD_ARG(p->numOfgradients >= 0, "para.numOfgradients must be >= 0");
    ASD_ARG(p->precision >= 0 && p->precision <= 5, "para.precision must be 0, 1, 2, 3, 4 or 5");
    return ASD_OK;
}

// Stage timing on the plan stream.  A one-kernel stage (a projector launch)
// gets its events from the dispatch itself (klaunch), so it measures the
// kernel's execution alone; a multi-kernel stage (a TV descent) is bracketed
// by events recorded before and after it.
struct StageTimer {
    asd_plan *p;
    int stage;
    bool one_kernel;
    hipEvent_t a = nullptr, b = nullptr;
    LaunchTiming lt{};
    StageTimer(asd_plan *pp, int st, bool one) : p(pp), stage(st), one_kernel(one)
    {
        if (!p->timing || p->capturing) return;
        if (hipEventCreate(&a) != hipSuccess) { a = nullptr; return; }
        if (hipEventCreate(&b) != hipSuccess) { (void)hipEventDestroy(a); a = nullptr; return; }
        if (one_kernel) {
            lt = LaunchTiming{a, b, 0};
            t_launch_timing = &lt;
        } else {
            (void)hipEventRecord(a, p->stream);
        }
    }
    ~StageTimer()
    {
        if (!a) return;
        if (one_kernel) {
            t_launch_timing = nullptr;
            if (lt.launches != 1) {  // nothing (or more than one kernel) was timed
                (void)hipEventDestroy(a);
                (void)hipEventDestroy(b);
                return;
            }
        } else {
            (void)hipEventRecord(b, p->stream);
        }
        p->ev.push_back({a, b, stage});
    }
};

void collect_timing(asd_plan *p)
{
    for (auto &e : p->ev) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
            p->stage_ms[e.stage] += ms;
            p->stage_cnt[e.stage] += 1;
        }
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    p->ev.clear();
}

// Codec of view iv of the plan's sinogram.
template <class Codec>
Codec view_codec(asd_plan *p, int iv);
template <>
SinoF32 view_codec<SinoF32>(asd_plan *p, int iv)
{
    return SinoF32{p->P.as<float>() + (size_t)iv * p->M};
}
template <>
SinoU16 view_codec<SinoU16>(asd_plan *p, int iv)
{
    return SinoU16{p->P.as<uint16_t>() + (size_t)iv * p->M, p->u16_lo[iv], p->u16_step[iv]};
}
template <>
SinoF16 view_codec<SinoF16>(asd_plan *p, int iv)
{
    return SinoF16{p->P.as<__half>() + (size_t)iv * p->M};
}
template <>
SinoF64 view_codec<SinoF64>(asd_plan *p, int iv)
{
    return SinoF64{p->P.as<double>() + (size_t)iv * p->M};
}

// One SART sweep over all views with relaxation *relax.  The precision
// variants read a 16-bit sinogram and compensate every FP32 sum; EXACT
// also sets up the backprojector's rays in FP64 per tile.
template <class Codec, bool KAHAN, bool EXACT>
int enqueue_sweep_t(asd_plan *p, const double *relax)
{
    const int *stopped = &p->ctl()->stopped;
    for (int iv = 0; iv < p->nv; iv++) {
        const ViewGeom &g = p->views[iv];
        const float *yd = p->ydet.as<float>() + (size_t)iv * p->para.na;
        const float *zd = p->zdet.as<float>() + (size_t)iv * p->para.nb;
        const RayRec *rt = p->rt.as<RayRec>() + (size_t)iv * p->M;
        EpiSartAx<Codec> ea{view_codec<Codec>(p, iv), p->cor.as<float>(), p->res2.as<double>(),
                            stopped, g.scale};
        int rc;
        {
            StageTimer t(p, ST_AX, true);
            // null detector tables: the gateway's grid is computed in the kernel
            if ((rc = dispatch_ax<KAHAN>(p->ax, g, p->m.as<float>(), nullptr, nullptr, ea,
                                         p->stream)))
                return rc;
        }
        {
            StageTimer t(p, ST_ATX, true);
            const EpiSartAtx eb{p->m.as<float>(), relax, stopped, g.scale};
            if (EXACT)
                rc = dispatch_atx_prec<KAHAN>(g, p->cor.as<float>(), yd, zd, rt, eb, p->stream);
            else
                rc = dispatch_atx_variant<KAHAN>(p->atx_variant, g, p->cor.as<float>(), rt, eb,
                                                 p->stream);
            if (rc) return rc;
        }
    }
    return ASD_OK;
}

// The sweep of the reference-arithmetic plan: the FP64 sinogram and image,
// the forward projector reading single(m) and summing every ray in slab
// order (as the reference kernel does), the backprojector's update in FP64.
// Every per-view quantity is the reference's arithmetic, so the sweep
// follows the reference bit for bit.  A forward shape forced with
// asd_plan_set_ax_shape replaces the slab-order projector (for measurement).
int enqueue_sweep_ref(asd_plan *p, const double *relax)
{
    const int *stopped = &p->ctl()->stopped;
    for (int iv = 0; iv < p->nv; iv++) {
        const ViewGeom &g = p->views[iv];
        const RayRec *rt = p->rt.as<RayRec>() + (size_t)iv * p->M;
        EpiSartAx<SinoF64> ea{view_codec<SinoF64>(p, iv), p->cor.as<float>(),
                              p->res2.as<double>(), stopped, g.scale};
        int rc;
        {
            StageTimer t(p, ST_AX, true);
            rc = p->ax_forced
                     ? dispatch_ax<false>(p->ax, g, p->ms.as<float>(), nullptr, nullptr, ea,
                                          p->stream)
                     : dispatch_ax_ordered(g, p->ms.as<float>(), nullptr, nullptr, ea, p->stream);
            if (rc) return rc;
        }
        {
            StageTimer t(p, ST_ATX, true);
            const EpiSartAtx64 eb{p->m.as<double>(), p->ms.as<float>(), relax, stopped, g.scale};
            if ((rc = dispatch_atx_variant<false>(p->atx_variant, g, p->cor.as<float>(), rt, eb,
                                                  p->stream)))
                return rc;
        }
    }
    return ASD_OK;
}

// Backprojector epilogue for the autotuner (plain x and sc).
struct EpiTune {
    float *x, *sc;
    float scale;
    struct Pre {};
    __device__ __forceinline__ bool skip() const { return false; }
    __device__ __forceinline__ Pre load(long) const { return Pre{}; }
    __device__ __forceinline__ void operator()(long v, float acc, float accl, const Pre &) const
    {
        x[v] = acc * scale;
        sc[v] = accl * scale;
    }
};

// Times the bit-identical backprojector variants on three views (the first,
// and those an eighth and a quarter of the way round; scratch outputs, before
// the first sweep) and keeps the fastest on average.  The choice only changes
// speed, never results.
template <bool KAHAN>
int autotune_atx_t(asd_plan *p);

int autotune_atx(asd_plan *p)
{
    if (p->atx_variant >= 0) return ASD_OK;
    for (int i = 0; i < ATX_NVARIANT; i++) p->atx_us[i] = -1.0;
    switch (p->para.precision) {
    case 0:
    case 4:
    case 5: return autotune_atx_t<false>(p);
    case 3: p->atx_variant = ATX_DIRECT; return ASD_OK;  // own tiled FP64-geometry kernel
    default: return autotune_atx_t<true>(p);
    }
}

template <bool KAHAN>
int autotune_atx_t(asd_plan *p)
{
    // the views timed: the first, and those an eighth and a quarter of the
    // way round (a variant's speed varies with the angle; one view misled
    // near-ties between sorted tiles)
    std::vector<int> tv = {0};
    for (int v : {p->nv / 8, p->nv / 4})
        if (v > tv.back()) tv.push_back(v);
    const ViewGeom &g = p->views[0];
    int rc = ASD_OK;
    ASD_HIP(hipMemsetAsync(p->cor.p, 0, p->M * 4, p->stream));
    std::vector<int> cand = {ATX_DIRECT, ATX_DIRECT2, ATX_DIRECT1, ATX_T16x16x1, ATX_PAIRS};
    if (g.nz > 1) {
        cand.push_back(ATX_T8x8x4);
        cand.push_back(ATX_T16x4x4);
        cand.push_back(ATX_SORTED16);
        cand.push_back(ATX_SORTED8);
        cand.push_back(ATX_SORTED8Z);
    }
    hipEvent_t e0, e1;
    ASD_HIP(hipEventCreate(&e0));
    ASD_HIP(hipEventCreate(&e1));
    const EpiTune ep{p->v.as<float>(), p->u2.as<float>(), g.scale};
    int best = ATX_DIRECT;
    double best_us = 1e30;
    for (int v : cand) {
        float ms_sum = 0.0f;
        for (int iv : tv) {
            float ms_min = 1e30f;
            for (int rep = 0; rep < 4; rep++) {
                // the kernel's own execution time (events stamped by the dispatch)
                LaunchTiming lt{e0, e1, 0};
                t_launch_timing = &lt;
                rc = dispatch_atx_variant<KAHAN>(v, p->views[iv], p->cor.as<float>(),
                                                 p->rt.as<RayRec>() + (size_t)iv * p->M, ep,
                                                 p->stream);
                t_launch_timing = nullptr;
                if (rc) break;
                (void)hipEventSynchronize(e1);
                float ms = 0;
                if (rep > 0 && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms < ms_min)
                    ms_min = ms;
            }
            if (rc) break;
            ms_sum += ms_min;
        }
        if (rc) break;
        const float ms_min = ms_sum / (float)tv.size();  // mean over the timed views
        p->atx_us[v] = ms_min * 1e3;
        // a later candidate must be faster by more than 3%: near-ties keep the
        // earlier one, so the choice does not flip between runs on timing noise
        if (ms_min * 1e3 < best_us * 0.97) {
            best_us = ms_min * 1e3;
            best = v;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    p->atx_variant = best;
    return ASD_OK;
}

int enqueue_sweep(asd_plan *p, const double *relax)
{
    switch (p->para.precision) {
    case 1: return enqueue_sweep_t<SinoU16, true, false>(p, relax);
    case 2: return enqueue_sweep_t<SinoF16, true, false>(p, relax);
    case 5: return enqueue_sweep_t<SinoF16, false, false>(p, relax);
    case 3: return enqueue_sweep_t<SinoU16, true, true>(p, relax);
    case 4: return enqueue_sweep_ref(p, relax);
    default: return enqueue_sweep_t<SinoF32, false, false>(p, relax);
    }
}

template <class T>
int enqueue_post_sart_t(asd_plan *p, int sweep_only)
{
    const unsigned nb = red_blocks(p->N > p->M ? p->N : p->M);
    post_sart_kernel<T><<<nb, RED_THREADS, 0, p->stream>>>(
        p->m.as<T>(), p->mpre.as<T>(), p->mk1.as<T>(), p->have_I0 ? p->I0.as<T>() : nullptr,
        p->f0.as<T>(), p->res2.as<double>(), p->N, p->M, p->red.partials, p->ctl());
    ASD_LAUNCH_CHECK("post_sart_kernel");
    // the reference-arithmetic plan sums dp^2, d^2 and rmse^2 in index order,
    // as the reference's loops (and the oracle) do
    const bool serial = p->ref64();
    if (serial) {
        const int *stopped = &p->ctl()->stopped;
        double *sc = p->red.scalars;
        int rc;
        if ((rc = launch_serial_sumsq<T>(p->m.as<T>(), p->mpre.as<T>(), p->N, sc + TV_SLOT_DP2,
                                         p->stream, stopped)) ||
            (rc = launch_serial_sumsq<T>(p->m.as<T>(), p->mk1.as<T>(), p->N, sc + TV_SLOT_D2,
                                         p->stream, stopped)) ||
            (p->have_I0 && (rc = launch_serial_sumsq<T>(p->m.as<T>(), p->I0.as<T>(), p->N,
                                                        sc + TV_SLOT_RMSE2, p->stream, stopped))))
            return rc;
    }
    ctl_after_sart<<<1, RED_THREADS, 0, p->stream>>>(p->red.partials, (int)nb, p->N,
                                                     p->have_I0 ? 1 : 0, p->red.scalars, p->ctl(),
                                                     p->hist.as<double>(), p->rules(), sweep_only,
                                                     serial ? 1 : 0);
    ASD_LAUNCH_CHECK("ctl_after_sart");
    return ASD_OK;
}

int enqueue_post_sart(asd_plan *p, int sweep_only)
{
    return p->ref64() ? enqueue_post_sart_t<double>(p, sweep_only)
                      : enqueue_post_sart_t<float>(p, sweep_only);
}

// TV descent of the loop (ASD_POCS_cuda.m:155-164) on the image type T, and
// post_tv (dg, m_pre = m, and the shadow of the reference-arithmetic plan).
template <class T>
int enqueue_tv_post(asd_plan *p)
{
    int rc;
    DevCtl *ctl = p->ctl();
    const asd_para &pa = p->para;
    const bool three_d = pa.nz >= 2;
    {
        StageTimer t(p, ST_TVGRAD, false);
        if ((rc = tv_descent<T, T>(p->m.as<T>(), p->mk1.as<T>(), p->v.as<T>(), p->u2.as<T>(),
                                   p->v2.as<T>(), pa.nx, pa.ny, three_d ? pa.nz : 1, three_d, 0.0,
                                   &ctl->alpha, pa.numOfgradients, 1, p->red, p->stream, true,
                                   &ctl->stopped, p->ref64())))
            return rc;
    }
    const unsigned nb = red_blocks(p->N);
    post_tv_kernel<T><<<nb, RED_THREADS, 0, p->stream>>>(
        p->m.as<T>(), p->f0.as<T>(), p->mpre.as<T>(), p->ref64() ? p->ms.as<float>() : nullptr,
        p->N, p->red.partials, ctl);
    ASD_LAUNCH_CHECK("post_tv_kernel");
    if (p->ref64())
        return launch_serial_sumsq<T>(p->m.as<T>(), p->f0.as<T>(), p->N,
                                      p->red.scalars + TV_SLOT_DG2, p->stream, &ctl->stopped);
    return ASD_OK;
}

// The kernels of one outer iteration (no host synchronisation inside).
int enqueue_iteration(asd_plan *p)
{
    int rc;
    DevCtl *ctl = p->ctl();
    if ((rc = enqueue_sweep(p, &ctl->beta))) return rc;
    if ((rc = enqueue_post_sart(p, 0))) return rc;
    if ((rc = p->ref64() ? enqueue_tv_post<double>(p) : enqueue_tv_post<float>(p))) return rc;
    const unsigned nb = red_blocks(p->N);
    ctl_after_tv<<<1, RED_THREADS, 0, p->stream>>>(p->red.partials, (int)nb, p->red.scalars, ctl,
                                                   p->rules(), p->ref64() ? 1 : 0);
    ASD_LAUNCH_CHECK("ctl_after_tv");
    return ASD_OK;
}

// An executable graph must not be destroyed while a replay of it may still
// run: the stream is drained first (every caller has drained it already; the
// drain keeps that true for any caller).
void drop_graph(asd_plan *p)
{
    if (p->exec && p->stream) (void)hipStreamSynchronize(p->stream);
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
}

int build_graph(asd_plan *p)
{
    ASD_HIP(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
    p->capturing = true;
    int rc = enqueue_iteration(p);
    p->capturing = false;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(p->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) {
        set_error(ASD_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        return ASD_ERR_HIP;
    }
    hipError_t ei = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        p->exec = nullptr;
        set_error(ASD_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        return ASD_ERR_HIP;
    }
    p->graph = g;
    return ASD_OK;
}

int sync_ctl(asd_plan *p)
{
    ASD_HIP(hipMemcpyAsync(&p->h, p->ctlbuf.p, sizeof(DevCtl), hipMemcpyDeviceToHost, p->stream));
    ASD_HIP(hipStreamSynchronize(p->stream));
    if (p->timing) collect_timing(p);
    return ASD_OK;
}

int write_ctl(asd_plan *p)
{
    ASD_HIP(hipMemcpyAsync(p->ctlbuf.p, &p->h, sizeof(DevCtl), hipMemcpyHostToDevice, p->stream));
    ASD_HIP(hipStreamSynchronize(p->stream));
    return ASD_OK;
}

// m_pre = m_k1 = f0 = m: the current image becomes the start of the loop
// (m_k1 is m_pre of the first iteration, ASD_POCS_cuda.m:151-153).
int rebase(asd_plan *p)
{
    if (p->ref64())
        copy3_kernel<double><<<grid_for(p->N), 256, 0, p->stream>>>(
            p->m.as<double>(), p->mpre.as<double>(), p->mk1.as<double>(), p->f0.as<double>(), p->N);
    else
        copy3_kernel<float><<<grid_for(p->N), 256, 0, p->stream>>>(
            p->m.as<float>(), p->mpre.as<float>(), p->mk1.as<float>(), p->f0.as<float>(), p->N);
    ASD_LAUNCH_CHECK("copy3_kernel");
    ASD_HIP(hipStreamSynchronize(p->stream));
    return ASD_OK;
}

int upload_f64_as_f32(asd_plan *p, DeviceBuf &dst, const double *src, size_t n)
{
    DeviceBuf tmp;
    int rc;
    if ((rc = tmp.ensure(n * sizeof(double)))) return rc;
    ASD_HIP(hipMemcpyAsync(tmp.p, src, n * sizeof(double), hipMemcpyHostToDevice, p->stream));
    f64_to_f32<<<grid_for(n), 256, 0, p->stream>>>(tmp.as<double>(), dst.as<float>(), n);
    ASD_LAUNCH_CHECK("f64_to_f32");
    ASD_HIP(hipStreamSynchronize(p->stream));
    return ASD_OK;
}

// An image of the plan from the host's FP64: FP64 as it is in the
// reference-arithmetic plan (with `shadow`, when not NULL, receiving
// single()), FP32 otherwise.
int upload_image(asd_plan *p, DeviceBuf &dst, const double *src, size_t n, DeviceBuf *shadow)
{
    if (!p->ref64()) return upload_f64_as_f32(p, dst, src, n);
    ASD_HIP(hipMemcpyAsync(dst.p, src, n * sizeof(double), hipMemcpyHostToDevice, p->stream));
    if (shadow) {
        f64_to_f32<<<grid_for(n), 256, 0, p->stream>>>(dst.as<double>(), shadow->as<float>(), n);
        ASD_LAUNCH_CHECK("f64_to_f32");
    }
    ASD_HIP(hipStreamSynchronize(p->stream));
    return ASD_OK;
}

int download_image(asd_plan *p, const DeviceBuf &src, double *dst, size_t n);

int download_f32_as_f64(asd_plan *p, const DeviceBuf &src, double *dst, size_t n)
{
    std::vector<float> h(n);
    ASD_HIP(hipMemcpyAsync(h.data(), src.p, n * sizeof(float), hipMemcpyDeviceToHost, p->stream));
    ASD_HIP(hipStreamSynchronize(p->stream));
    for (size_t i = 0; i < n; i++) dst[i] = (double)h[i];
    return ASD_OK;
}

int download_image(asd_plan *p, const DeviceBuf &src, double *dst, size_t n)
{
    if (!p->ref64()) return download_f32_as_f64(p, src, dst, n);
    ASD_HIP(hipMemcpyAsync(dst, src.p, n * sizeof(double), hipMemcpyDeviceToHost, p->stream));
    ASD_HIP(hipStreamSynchronize(p->stream));
    return ASD_OK;
}

int need_projections(asd_plan *p)
{
    if (!p->have_P) {
        set_error(ASD_ERR_STATE, "plan has no projections (call asd_plan_set_projections)");
        return ASD_ERR_STATE;
    }
    return ASD_OK;
}

}  // namespace

void asd::plan_drop_graph(asd_plan *p) { drop_graph(p); }

extern "C" {

void asd_para_defaults(asd_para *p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->nz = 1;
    p->nb = 1;
    p->maxnumIter = 5;
    p->numOfgradients = 20;
    p->alpha = 0.02;
    p->alpha_red = 0.95;
    p->beta = 1.0;
    p->beta_red = 0.995;
    p->r_max = 0.95;
    p->epsilon = 2.220446049250313e-16;
    p->stopRelChg = 2.220446049250313e-16;
    p->initBySART = 0;
    p->lambda = 1.0;
    p->precision = 0;
}

int asd_plan_create(const asd_para *para, const double *phi_vec, int nv, asd_plan **out)
{
    clear_error();
    ASD_ARG(out, "NULL plan output");
    *out = nullptr;
    int rc = check_para(para, nv);
    if (rc) return rc;
    ASD_ARG(phi_vec, "NULL phi_vec");
    asd_plan *p = new asd_plan();
    p->para = *para;
    p->nv = nv;
    p->N = (size_t)para->nx * para->ny * para->nz;
    p->M = (size_t)para->na * para->nb;
    std::vector<float> yall, zall, yd, zd;
    p->views.resize(nv);
    for (int i = 0; i < nv; i++) {
        view_geom_from_para(*para, phi_vec[i], p->views[i], yd, zd);
        yall.insert(yall.end(), yd.begin(), yd.end());
        zall.insert(zall.end(), zd.begin(), zd.end());
    }
    const size_t N = p->N, M = p->M;
    p->ax = ax_shape(para->na, para->nb, para->nx > para->ny ? para->nx : para->ny);
    auto fail = [&](int code) { asd_plan_destroy(p); return code; };
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
        p->stream = nullptr;
        set_error(ASD_ERR_HIP, "hipStreamCreateWithFlags failed");
        return fail(ASD_ERR_HIP);
    }
    if (hipHostMalloc((void **)&p->h_stop, 4 * sizeof(int), hipHostMallocDefault) != hipSuccess) {
        p->h_stop = nullptr;
        set_error(ASD_ERR_HIP, "hipHostMalloc failed");
        return fail(ASD_ERR_HIP);
    }
    for (hipEvent_t &e : p->stop_ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            e = nullptr;
            set_error(ASD_ERR_HIP, "hipEventCreate failed");
            return fail(ASD_ERR_HIP);
        }
    const size_t ib = p->img_bytes();
    if ((rc = p->ydet.ensure(yall.size() * 4)) || (rc = p->zdet.ensure(zall.size() * 4)) ||
        (rc = p->m.ensure(N * ib)) || (rc = p->mpre.ensure(N * ib)) ||
        (rc = p->mk1.ensure(N * ib)) || (rc = p->f0.ensure(N * ib)) || (rc = p->v.ensure(N * ib)) ||
        (rc = p->u2.ensure(N * ib)) || (rc = p->v2.ensure(N * ib)) ||
        (p->ref64() && (rc = p->ms.ensure(N * 4))) ||
        (rc = p->cor.ensure(M * 4)) || (rc = p->res2.ensure(M * 8)) ||
        (rc = p->rt.ensure((size_t)nv * M * sizeof(RayRec))) ||
        (rc = p->P.ensure((size_t)nv * M * sino_bytes(para->precision))) ||
        (rc = p->ctlbuf.ensure(sizeof(DevCtl))) || (rc = p->hist.ensure(HIST_CAP * 8)) ||
        (rc = p->red.init()))
        return fail(rc);
    if (hipMemcpy(p->ydet.p, yall.data(), yall.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->zdet.p, zall.data(), zall.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(p->m.p, 0, N * ib) != hipSuccess || hipMemset(p->res2.p, 0, M * 8) != hipSuccess) {
        set_error(ASD_ERR_HIP, "plan initialisation failed");
        return fail(ASD_ERR_HIP);
    }
    // the ray-slope tables of every view, once: they depend on geometry only
    for (int i = 0; i < nv; i++)
        if ((rc = launch_ray_table(p->views[i], p->ydet.as<float>() + (size_t)i * para->na,
                                   p->zdet.as<float>() + (size_t)i * para->nb,
                                   p->rt.as<RayRec>() + (size_t)i * M, p->stream)))
            return fail(rc);
    *out = p;
    rc = asd_plan_reset(p, nullptr, nullptr);
    if (rc) {
        *out = nullptr;
        return fail(rc);
    }
    return ASD_OK;
}

int asd_plan_set_projections(asd_plan *p, const double *P)
{
    clear_error();
    ASD_ARG(p && P, "NULL argument");
    int rc;
    const size_t n = (size_t)p->nv * p->M;
    if (p->para.precision == 1 || p->para.precision == 3) {
        // 16-bit fixed point per view: lo + step * q, q = round((P - lo) / step)
        std::vector<uint16_t> q(n);
        p->u16_lo.assign(p->nv, 0.0);
        p->u16_step.assign(p->nv, 0.0);
        for (int iv = 0; iv < p->nv; iv++) {
            const double *v = P + (size_t)iv * p->M;
            double lo = v[0], hi = v[0];
            for (size_t i = 1; i < p->M; i++) {
                lo = v[i] < lo ? v[i] : lo;
                hi = v[i] > hi ? v[i] : hi;
            }
            const double step = hi > lo ? (hi - lo) / 65535.0 : 0.0;
            p->u16_lo[iv] = lo;
            p->u16_step[iv] = step;
            for (size_t i = 0; i < p->M; i++) {
                double t = step > 0 ? (v[i] - lo) / step : 0.0;
                long r = std::lround(t);
                q[(size_t)iv * p->M + i] = (uint16_t)(r < 0 ? 0 : r > 65535 ? 65535 : r);
            }
        }
        ASD_HIP(hipMemcpy(p->P.p, q.data(), n * 2, hipMemcpyHostToDevice));
        drop_graph(p);  // the view scales are kernel arguments
    } else if (p->para.precision == 2 || p->para.precision == 5) {
        std::vector<__half> h(n);
        for (size_t i = 0; i < n; i++) h[i] = __float2half_rn((float)P[i]);
        ASD_HIP(hipMemcpy(p->P.p, h.data(), n * 2, hipMemcpyHostToDevice));
    } else if (p->para.precision == 4) {
        ASD_HIP(hipMemcpy(p->P.p, P, n * 8, hipMemcpyHostToDevice));  // FP64, as MATLAB holds it
    } else if ((rc = upload_f64_as_f32(p, p->P, P, n))) {
        return rc;
    }
    p->have_P = true;
    return ASD_OK;
}

int asd_plan_reset(asd_plan *p, const double *m0, const double *I0)
{
    clear_error();
    ASD_ARG(p, "NULL plan");
    int rc;
    if (m0) {
        if ((rc = upload_image(p, p->m, m0, p->N, &p->ms))) return rc;
    } else {
        ASD_HIP(hipMemsetAsync(p->m.p, 0, p->N * p->img_bytes(), p->stream));
        if (p->ref64()) ASD_HIP(hipMemsetAsync(p->ms.p, 0, p->N * 4, p->stream));
    }
    bool had_I0 = p->have_I0;
    if (I0) {
        if ((rc = p->I0.ensure(p->N * p->img_bytes()))) return rc;
        if ((rc = upload_image(p, p->I0, I0, p->N, nullptr))) return rc;
        p->have_I0 = true;
        p->I0_sim = false;
    } else {
        p->have_I0 = p->I0_sim;  // a simulated phantom stays the reference
    }
    if (had_I0 != p->have_I0) drop_graph(p);  // the graph bakes in the I0 pointer
    ASD_HIP(hipMemsetAsync(p->res2.p, 0, p->M * 8, p->stream));
    if ((rc = rebase(p))) return rc;
    std::memset(&p->h, 0, sizeof(p->h));
    p->h.alpha = p->para.alpha;
    p->h.beta = p->para.beta;
    p->h.relax = p->para.lambda;
    return write_ctl(p);
}

int asd_plan_iterate(asd_plan *p, int n_iter)
{
    clear_error();
    ASD_ARG(p, "NULL plan");
    ASD_ARG(n_iter >= 0, "n_iter must be >= 0");
    int rc;
    if ((rc = need_projections(p))) return rc;
    if (p->h.stopped || n_iter == 0) return ASD_OK;
    if ((rc = autotune_atx(p))) return rc;
    const bool graph = p->use_graph && !p->timing;
    if (graph && !p->exec && (rc = build_graph(p))) return rc;
    // The stop flag is copied to pinned host memory after every STOP_EVERY-th
    // iteration; the host reads the copy two checkpoints later, so once the
    // loop stops at most ~3 * STOP_EVERY further iterations are enqueued (their
    // kernels change nothing) while the GPU never waits for the host.
    constexpr int STOP_EVERY = 4;
    for (int i = 0; i < n_iter; i++) {
        if (i >= 2 * STOP_EVERY && i % STOP_EVERY == 0) {
            const int c = i / STOP_EVERY - 2;
            ASD_HIP(hipEventSynchronize(p->stop_ev[c % 4]));
            if (p->h_stop[c % 4]) break;
        }
        if (graph)
            ASD_HIP(hipGraphLaunch(p->exec, p->stream));
        else if ((rc = enqueue_iteration(p)))
            return rc;
        if (i % STOP_EVERY == STOP_EVERY - 1) {
            const int c = i / STOP_EVERY;
            ASD_HIP(hipMemcpyAsync(&p->h_stop[c % 4], &p->ctl()->stopped, sizeof(int),
                                   hipMemcpyDeviceToHost, p->stream));
            ASD_HIP(hipEventRecord(p->stop_ev[c % 4], p->stream));
        }
    }
    return sync_ctl(p);
}

int asd_plan_sweep(asd_plan *p, double lambda, double *dataTolerance)
{
    clear_error();
    ASD_ARG(p, "NULL plan");
    int rc;
    if ((rc = need_projections(p))) return rc;
    ASD_ARG(!p->h.stopped, "plan has stopped; reset it first");
    if ((rc = autotune_atx(p))) return rc;
    p->h.relax = lambda;
    if ((rc = write_ctl(p))) return rc;
    if ((rc = enqueue_sweep(p, &p->ctl()->relax))) return rc;
    if ((rc = enqueue_post_sart(p, 1))) return rc;
    if ((rc = sync_ctl(p))) return rc;
    if (dataTolerance) *dataTolerance = p->h.dd2;
    return rebase(p);
}

int asd_plan_get_image(asd_plan *p, double *m_res)
{
    clear_error();
    ASD_ARG(p && m_res, "NULL argument");
    return download_image(p, p->f0, m_res, p->N);
}

int asd_plan_get_current(asd_plan *p, double *m)
{
    clear_error();
    ASD_ARG(p && m, "NULL argument");
    return download_image(p, p->m, m, p->N);
}

int asd_plan_get_out(asd_plan *p, asd_out *o, double *rmse)
{
    clear_error();
    ASD_ARG(p && o, "NULL argument");
    o->IterationNum = p->h.k;
    o->dataTolerance = p->h.dd2;
    o->beta = p->h.beta;
    o->alpha = p->h.alpha;
    if (rmse && p->have_I0 && p->h.k > 0) {
        int n = p->h.k < HIST_CAP ? p->h.k : HIST_CAP;
        ASD_HIP(hipMemcpy(rmse, p->hist.p, n * sizeof(double), hipMemcpyDeviceToHost));
    }
    return ASD_OK;
}

int asd_plan_sync(asd_plan *p)
{
    clear_error();
    ASD_ARG(p, "NULL plan");
    ASD_HIP(hipStreamSynchronize(p->stream));
    return ASD_OK;
}

int asd_plan_set_timing(asd_plan *p, int enable)
{
    clear_error();
    ASD_ARG(p, "NULL plan");
    p->timing = enable != 0;
    for (int i = 0; i < ST_N; i++) {
        p->stage_ms[i] = 0;
        p->stage_cnt[i] = 0;
    }
    return ASD_OK;
}

int asd_plan_set_graph(asd_plan *p, int enable)
{
    clear_error();
    ASD_ARG(p, "NULL plan");
    p->use_graph = enable != 0;
    if (!p->use_graph) drop_graph(p);
    return ASD_OK;
}

int asd_plan_get_timing(asd_plan *p, int stage, double *total_ms, long *count)
{
    clear_error();
    ASD_ARG(p && total_ms && count, "NULL argument");
    ASD_ARG(stage >= 0 && stage < ST_N, "stage out of range");
    *total_ms = p->stage_ms[stage];
    *count = p->stage_cnt[stage];
    return ASD_OK;
}

int asd_plan_set_atx_variant(asd_plan *p, int atx_variant)
{
    clear_error();
    ASD_ARG(p, "NULL plan");
    ASD_ARG(atx_variant >= -1 && atx_variant < ATX_NVARIANT, "atx_variant must be -1..9");
    ASD_ARG((atx_variant != ATX_T8x8x4 && atx_variant != ATX_T16x4x4 &&
             atx_variant != ATX_SORTED8 && atx_variant != ATX_SORTED8Z) || p->views[0].nz > 1,
            "the 3-D tile variants need nz > 1");
    drop_graph(p);
    p->atx_variant = atx_variant;
    for (int i = 0; i < ATX_NVARIANT; i++) p->atx_us[i] = -1.0;
    return ASD_OK;
}

int asd_plan_set_ax_shape(asd_plan *p, int lanes_per_ray, int rows_per_bundle)
{
    clear_error();
    ASD_ARG(p, "NULL plan");
    if (lanes_per_ray == 0 && rows_per_bundle == 0) {
        p->ax = ax_shape(p->para.na, p->para.nb, std::max(p->para.nx, p->para.ny));
        p->ax_forced = false;
    } else {
        const int l = lanes_per_ray, r = rows_per_bundle;
        ASD_ARG((r == 1 && l >= 1 && l <= 256 && (l & (l - 1)) == 0) ||
                    ((r == 2 || r == 3) && l == 8),
                "ax shape must be (1..256 lanes, a power of two; 1 row) or (8 lanes; 2 or 3 "
                "rows), or (0, 0) for the default");
        p->ax.lpr = l;
        p->ax.rows = r;
        p->ax_forced = true;
    }
    drop_graph(p);
    return ASD_OK;
}

int asd_plan_tuning(asd_plan *p, int *atx_variant, double *atx_us)
{
    clear_error();
    ASD_ARG(p, "NULL plan");
    if (atx_variant) *atx_variant = p->atx_variant;
    if (atx_us)
        for (int i = 0; i < ATX_NVARIANT; i++) atx_us[i] = p->atx_variant >= 0 ? p->atx_us[i] : -1.0;
    return ASD_OK;
}

int asd_plan_info(asd_plan *p, int *lanes_per_ray, int *rows_per_bundle, void **stream)
{
    clear_error();
    ASD_ARG(p, "NULL plan");
    // the reference-arithmetic plan's slab-order projector reports 0 lanes
    const bool ordered = p->ref64() && !p->ax_forced;
    if (lanes_per_ray) *lanes_per_ray = ordered ? 0 : p->ax.lpr;
    if (rows_per_bundle) *rows_per_bundle = ordered ? 1 : p->ax.rows;
    if (stream) *stream = (void *)p->stream;
    return ASD_OK;
}

void asd_plan_destroy(asd_plan *p)
{
    if (!p) return;
    if (p->stream) {
        (void)hipStreamSynchronize(p->stream);
        collect_timing(p);
    }
    drop_graph(p);
    for (hipEvent_t &e : p->stop_ev)
        if (e) (void)hipEventDestroy(e);
    if (p->h_stop) (void)hipHostFree(p->h_stop);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

int asd_pocs(const double *P, const double *phi_vec, int nv, const asd_para *para,
             const double *m0, const double *I0, double *m_res, double *rmse, asd_out *out)
{
    clear_error();
    ASD_ARG(P && phi_vec && m_res && out, "NULL argument");
    asd_plan *p = nullptr;
    int rc = asd_plan_create(para, phi_vec, nv, &p);
    if (rc) return rc;
    auto done = [&](int code) {
        asd_plan_destroy(p);
        return code;
    };
    if ((rc = asd_plan_set_projections(p, P))) return done(rc);
    // m0: para.m0, else one SART sweep when initBySART, else zeros
    // (ASD_POCS_cuda.m:20-32).
    if ((rc = asd_plan_reset(p, m0, I0))) return done(rc);
    if (!m0 && para->initBySART) {
        if ((rc = asd_plan_sweep(p, para->lambda, nullptr))) return done(rc);
    }
    if ((rc = asd_plan_iterate(p, para->maxnumIter))) return done(rc);
    if ((rc = asd_plan_get_image(p, m_res))) return done(rc);
    if ((rc = asd_plan_get_out(p, out, rmse))) return done(rc);
    return done(ASD_OK);
}

int asd_sart(const double *P, const double *phi_vec, int nv, const asd_para *para, double *m,
             double *dataTolerance)
{
    clear_error();
    ASD_ARG(P && phi_vec && m, "NULL argument");
    asd_plan *p = nullptr;
    int rc = asd_plan_create(para, phi_vec, nv, &p);
    if (rc) return rc;
    auto done = [&](int code) {
        asd_plan_destroy(p);
        return code;
    };
    if ((rc = asd_plan_set_projections(p, P))) return done(rc);
    if ((rc = asd_plan_reset(p, m, nullptr))) return done(rc);
    double tol = 0;
    if ((rc = asd_plan_sweep(p, para->lambda, &tol))) return done(rc);
    if ((rc = asd_plan_get_current(p, m))) return done(rc);
    if (dataTolerance) *dataTolerance = tol;
    return done(ASD_OK);
}

}  // extern "C"
