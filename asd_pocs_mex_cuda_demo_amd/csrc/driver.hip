// Device-resident ASD-POCS driver: the loop of utilities/ASD_POCS_cuda.m:103-171
// and the sweep of SART_cuda_inline.m:56-87 with the image, the sinogram and
// every per-view quantity kept in HBM.  Per view two kernels run:
//   ax_sart   ray sums -> cor = (P - vxl*y) / (vxl*sr | 1), the data-error term
//             and the view's ray-slope table (ASD_POCS_cuda.m:117-123)
//   atx_sart  voxel sums -> m = max(0, m + beta * (vxl*x) / (vxl*sc | 1))
//             (ASD_POCS_cuda.m:125-131)
// and per outer iteration FP64 grid reductions give dd2, dp, d, dg and the
// RMSE; the host reads them once per iteration to apply the alpha/beta rules.
#include <cmath>
#include <cstring>
#include <vector>

#include "projector_kernels.hpp"
#include "reduce_device.hpp"
#include "tv.hpp"

namespace asd {

namespace {

// Stage timing (events on the plan stream) for the dominant kernels.
enum : int { ST_AX = 0, ST_ATX = 1, ST_TVGRAD = 2, ST_N = 3 };

template <class PT>
struct EpiSartAx {
    const PT *P;          // this view's measured projection
    float *cor;           // correction fed to the backprojector
    double *res2;         // per-ray accumulated squared residual of the sweep
    RayRec *rt;           // ray-slope table of this view (for atx_sart)
    const float *y_det, *z_det;
    ViewGeom g;
    __device__ __forceinline__ void operator()(long id, float acc, float accl) const
    {
#pragma clang fp contract(off)
        float proj = acc * g.scale, sumr = accl * g.scale;
        double r = (sumr == 0.0f) ? 1.0 : (double)sumr;
        double c0 = (double)(float)P[id] - (double)proj;
        res2[id] += c0 * c0;
        cor[id] = (float)(c0 / r);
        int ia = (int)(id % g.na), ib = (int)(id / g.na);
        rt[id] = make_ray_rec(g, y_det[ia], z_det[ib]);
    }
};

struct EpiSartAtx {
    float *m;
    double beta;
    float scale;
    __device__ __forceinline__ void operator()(long v, float acc, float accl) const
    {
#pragma clang fp contract(off)
        float bp = acc * scale, sc = accl * scale;
        double c = (sc == 0.0f) ? 1.0 : (double)sc;
        double nm = (double)m[v] + beta * ((double)bp / c);
        m[v] = nm < 0 ? 0.0f : (float)nm;
    }
};

__global__ __launch_bounds__(256) void fill_f64(double *p, size_t n, double val)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = val;
}

__global__ __launch_bounds__(256) void sum_f64(const double *__restrict__ a, size_t n,
                                               double *partials, unsigned *counter, double *out)
{
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        s += a[i];
    grid_sum_to(s, partials, counter, out);
}

__global__ __launch_bounds__(256) void f64_to_f32(const double *__restrict__ a, float *b,
                                                  size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
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

struct asd_plan {
    asd_para para;
    int nv = 0;
    size_t N = 0, M = 0;
    std::vector<ViewGeom> views;
    hipStream_t stream = nullptr;
    DeviceBuf ydet, zdet, P, m, mpre, mk1, f0, I0, cor, res2, rt, v, stage;
    TvScratch red;
    bool have_P = false, have_I0 = false, stopped = false;
    int k = 0;
    double alpha = 0, beta = 0, dd2 = 0;
    std::vector<double> rmse;
    int lpr = 1;
    // optional stage timing
    bool timing = false;
    struct Ev { hipEvent_t a, b; int stage; };
    std::vector<Ev> ev;
    double stage_ms[ST_N] = {0, 0, 0};
    long stage_cnt[ST_N] = {0, 0, 0};
};

namespace {

int view_geom_from_para(const asd_para &p, double phi_deg, ViewGeom &g, std::vector<float> &ydet,
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
    ydet.resize(p.na);
    zdet.resize(p.nb);
    for (int i = 0; i < p.na; i++)
        ydet[i] = (float)(((double)(float)i - (double)(float)p.na / 2.0 + 0.5) * (double)det);
    for (int j = 0; j < p.nb; j++)
        zdet[j] = (float)(((double)(float)j - (double)(float)p.nb / 2.0 + 0.5) * (double)det);
    return ASD_OK;
}

int check_para(const asd_para *p, int nv)
{
    ASD_ARG(p, "NULL para");
    ASD_ARG(p->nx > 0 && p->ny > 0 && p->nz > 0, "para.nx, para.ny, para.nz must be positive");
    ASD_ARG(p->na > 0 && p->nb > 0, "para.na, para.nb must be positive");
    ASD_ARG(nv > 0, "phi_vec is empty");
    ASD_ARG(p->vxlsize > 0 && p->detsize > 0, "para.vxlsize and para.detsize must be positive");
    ASD_ARG(p->SO > 0 && p->OD > 0, "para.SO and para.OD must be positive");
    ASD_ARG(p->precision == 0 || p->precision == 1, "para.precision must be 0 or 1");
    return ASD_OK;
}

// Timed launch helper: events bracket the kernel on the plan stream.
struct StageTimer {
    asd_plan *p;
    int stage;
    hipEvent_t a = nullptr, b = nullptr;
    StageTimer(asd_plan *pp, int st) : p(pp), stage(st)
    {
        if (!p->timing) return;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
            a = b = nullptr;
            return;
        }
        hipEventRecord(a, p->stream);
    }
    ~StageTimer()
    {
        if (!a) return;
        hipEventRecord(b, p->stream);
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

template <class PT>
int sweep(asd_plan *p, double beta)
{
    const size_t M = p->M;
    fill_f64<<<grid_for(M), 256, 0, p->stream>>>(p->res2.as<double>(), M, 0.0);
    ASD_LAUNCH_CHECK("fill_f64");
    const long N = (long)p->N;
    const unsigned atx_blocks = (unsigned)((N + 255) / 256);
    for (int iv = 0; iv < p->nv; iv++) {
        const ViewGeom &g = p->views[iv];
        const float *yd = p->ydet.as<float>() + (size_t)iv * p->para.na;
        const float *zd = p->zdet.as<float>() + (size_t)iv * p->para.nb;
        EpiSartAx<PT> ea{p->P.as<PT>() + (size_t)iv * M, p->cor.as<float>(), p->res2.as<double>(),
                         p->rt.as<RayRec>(), yd, zd, g};
        int rc;
        {
            StageTimer t(p, ST_AX);
            if ((rc = dispatch_ax(p->lpr, g, p->m.as<float>(), yd, zd, ea, p->stream))) return rc;
        }
        {
            StageTimer t(p, ST_ATX);
            atx_kernel<<<atx_blocks, 256, 0, p->stream>>>(
                g, p->cor.as<float>(), p->rt.as<RayRec>(),
                EpiSartAtx{p->m.as<float>(), beta, g.scale});
            ASD_LAUNCH_CHECK("atx_kernel(sart)");
        }
    }
    // data error of the sweep
    sum_f64<<<red_blocks(M), RED_THREADS, 0, p->stream>>>(
        p->res2.as<double>(), M, p->red.partials, p->red.counters + TV_SLOT_DD2,
        p->red.scalars + TV_SLOT_DD2);
    ASD_LAUNCH_CHECK("sum_f64");
    return ASD_OK;
}

int run_sweep(asd_plan *p, double beta)
{
    return sweep<float>(p, beta);
}

int read_scalars(asd_plan *p, double *h)
{
    ASD_HIP(hipMemcpyAsync(h, p->red.scalars, sizeof(double) * TV_NSLOTS, hipMemcpyDeviceToHost,
                           p->stream));
    ASD_HIP(hipStreamSynchronize(p->stream));
    if (p->timing) collect_timing(p);
    return ASD_OK;
}

// One outer iteration k = p->k + 1 (ASD_POCS_cuda.m:103-171).
int iterate_once(asd_plan *p)
{
    const int k = p->k + 1;
    const size_t N = p->N;
    const size_t NB = N * sizeof(float);
    int rc;
    ASD_HIP(hipMemcpyAsync(p->mpre.p, p->m.p, NB, hipMemcpyDeviceToDevice, p->stream));
    if (k == 1) ASD_HIP(hipMemcpyAsync(p->mk1.p, p->m.p, NB, hipMemcpyDeviceToDevice, p->stream));
    if ((rc = run_sweep(p, p->beta))) return rc;
    float *m = p->m.as<float>();
    if ((rc = launch_sqdiff<float>(m, p->mpre.as<float>(), N, p->red, TV_SLOT_DP2, p->stream)))
        return rc;
    if ((rc = launch_sqdiff<float>(m, p->mk1.as<float>(), N, p->red, TV_SLOT_D2, p->stream)))
        return rc;
    if (p->have_I0 &&
        (rc = launch_sqdiff<float>(m, p->I0.as<float>(), N, p->red, TV_SLOT_RMSE2, p->stream)))
        return rc;
    ASD_HIP(hipMemcpyAsync(p->f0.p, p->m.p, NB, hipMemcpyDeviceToDevice, p->stream));
    // TV descent from m_k1 (the image before the first sweep), then dg.
    const asd_para &pa = p->para;
    const bool three_d = pa.nz >= 2;
    if ((rc = tv_descent<float, float>(m, p->mk1.as<float>(), p->v.as<float>(), pa.nx, pa.ny,
                                       three_d ? pa.nz : 1, three_d, p->alpha, nullptr,
                                       pa.numOfgradients, 1, p->red, p->stream, true)))
        return rc;
    if ((rc = launch_sqdiff<float>(m, p->f0.as<float>(), N, p->red, TV_SLOT_DG2, p->stream)))
        return rc;
    double h[TV_NSLOTS];
    if ((rc = read_scalars(p, h))) return rc;
    p->k = k;
    p->dd2 = h[TV_SLOT_DD2];
    const double dd = std::sqrt(h[TV_SLOT_DD2]);
    if (p->have_I0) p->rmse.push_back(std::sqrt(h[TV_SLOT_RMSE2] / (double)N));
    const double dp = std::sqrt(h[TV_SLOT_DP2]);
    if (dp / (double)N <= pa.stopRelChg) {
        // m_res is the post-SART image (f0); the TV result is discarded.
        p->stopped = true;
        return ASD_OK;
    }
    const double dg = std::sqrt(h[TV_SLOT_DG2]);
    if (dg > pa.r_max * dp && dd > pa.epsilon) p->alpha = p->alpha * pa.alpha_red;
    p->beta = p->beta * pa.beta_red;
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

int download_f32_as_f64(asd_plan *p, const DeviceBuf &src, double *dst, size_t n)
{
    std::vector<float> h(n);
    ASD_HIP(hipMemcpyAsync(h.data(), src.p, n * sizeof(float), hipMemcpyDeviceToHost, p->stream));
    ASD_HIP(hipStreamSynchronize(p->stream));
    for (size_t i = 0; i < n; i++) dst[i] = (double)h[i];
    return ASD_OK;
}

}  // namespace

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
    p->lpr = ax_lanes_per_ray(para->na, para->nb, para->nx > para->ny ? para->nx : para->ny);
    auto fail = [&](int code) { asd_plan_destroy(p); return code; };
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error(ASD_ERR_HIP, "hipStreamCreateWithFlags failed");
        return fail(ASD_ERR_HIP);
    }
    if ((rc = p->ydet.ensure(yall.size() * 4)) || (rc = p->zdet.ensure(zall.size() * 4)) ||
        (rc = p->m.ensure(N * 4)) || (rc = p->mpre.ensure(N * 4)) ||
        (rc = p->mk1.ensure(N * 4)) || (rc = p->f0.ensure(N * 4)) || (rc = p->v.ensure(N * 4)) ||
        (rc = p->cor.ensure(M * 4)) || (rc = p->res2.ensure(M * 8)) ||
        (rc = p->rt.ensure(M * sizeof(RayRec))) || (rc = p->P.ensure((size_t)nv * M * 4)) ||
        (rc = p->red.init()))
        return fail(rc);
    if (hipMemcpy(p->ydet.p, yall.data(), yall.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->zdet.p, zall.data(), zall.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(p->m.p, 0, N * 4) != hipSuccess) {
        set_error(ASD_ERR_HIP, "geometry upload failed");
        return fail(ASD_ERR_HIP);
    }
    p->alpha = para->alpha;
    p->beta = para->beta;
    *out = p;
    return ASD_OK;
}

int asd_plan_set_projections(asd_plan *p, const double *P)
{
    clear_error();
    ASD_ARG(p && P, "NULL argument");
    int rc = upload_f64_as_f32(p, p->P, P, (size_t)p->nv * p->M);
    if (rc) return rc;
    p->have_P = true;
    return ASD_OK;
}

int asd_plan_reset(asd_plan *p, const double *m0, const double *I0)
{
    clear_error();
    ASD_ARG(p, "NULL plan");
    int rc;
    if (m0) {
        if ((rc = upload_f64_as_f32(p, p->m, m0, p->N))) return rc;
    } else {
        ASD_HIP(hipMemsetAsync(p->m.p, 0, p->N * 4, p->stream));
    }
    if (I0) {
        if ((rc = p->I0.ensure(p->N * 4))) return rc;
        if ((rc = upload_f64_as_f32(p, p->I0, I0, p->N))) return rc;
        p->have_I0 = true;
    } else {
        p->have_I0 = false;
    }
    ASD_HIP(hipMemcpyAsync(p->f0.p, p->m.p, p->N * 4, hipMemcpyDeviceToDevice, p->stream));
    ASD_HIP(hipStreamSynchronize(p->stream));
    p->k = 0;
    p->stopped = false;
    p->alpha = p->para.alpha;
    p->beta = p->para.beta;
    p->dd2 = 0;
    p->rmse.clear();
    return ASD_OK;
}

int asd_plan_iterate(asd_plan *p, int n_iter)
{
    clear_error();
    ASD_ARG(p, "NULL plan");
    if (!p->have_P) {
        set_error(ASD_ERR_STATE, "plan has no projections (call asd_plan_set_projections)");
        return ASD_ERR_STATE;
    }
    for (int i = 0; i < n_iter && !p->stopped; i++) {
        int rc = iterate_once(p);
        if (rc) return rc;
    }
    return ASD_OK;
}

int asd_plan_sweep(asd_plan *p, double lambda, double *dataTolerance)
{
    clear_error();
    ASD_ARG(p, "NULL plan");
    if (!p->have_P) {
        set_error(ASD_ERR_STATE, "plan has no projections (call asd_plan_set_projections)");
        return ASD_ERR_STATE;
    }
    int rc = run_sweep(p, lambda);
    if (rc) return rc;
    double h[TV_NSLOTS];
    if ((rc = read_scalars(p, h))) return rc;
    if (dataTolerance) *dataTolerance = h[TV_SLOT_DD2];
    ASD_HIP(hipMemcpyAsync(p->f0.p, p->m.p, p->N * 4, hipMemcpyDeviceToDevice, p->stream));
    ASD_HIP(hipStreamSynchronize(p->stream));
    return ASD_OK;
}

int asd_plan_get_image(asd_plan *p, double *m_res)
{
    clear_error();
    ASD_ARG(p && m_res, "NULL argument");
    return download_f32_as_f64(p, p->f0, m_res, p->N);
}

int asd_plan_get_current(asd_plan *p, double *m)
{
    clear_error();
    ASD_ARG(p && m, "NULL argument");
    return download_f32_as_f64(p, p->m, m, p->N);
}

int asd_plan_get_out(asd_plan *p, asd_out *o, double *rmse)
{
    clear_error();
    ASD_ARG(p && o, "NULL argument");
    o->IterationNum = p->k;
    o->dataTolerance = p->dd2;
    o->beta = p->beta;
    o->alpha = p->alpha;
    if (rmse)
        for (size_t i = 0; i < p->rmse.size(); i++) rmse[i] = p->rmse[i];
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

int asd_plan_get_timing(asd_plan *p, int stage, double *total_ms, long *count)
{
    clear_error();
    ASD_ARG(p && total_ms && count, "NULL argument");
    ASD_ARG(stage >= 0 && stage < ST_N, "stage out of range");
    *total_ms = p->stage_ms[stage];
    *count = p->stage_cnt[stage];
    return ASD_OK;
}

int asd_plan_info(asd_plan *p, int *lanes_per_ray, void **stream)
{
    clear_error();
    ASD_ARG(p, "NULL plan");
    if (lanes_per_ray) *lanes_per_ray = p->lpr;
    if (stream) *stream = (void *)p->stream;
    return ASD_OK;
}

void asd_plan_destroy(asd_plan *p)
{
    if (!p) return;
    if (p->stream) {
        (void)hipStreamSynchronize(p->stream);
        collect_timing(p);
        (void)hipStreamDestroy(p->stream);
    }
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
        double tol;
        if ((rc = asd_plan_sweep(p, para->lambda, &tol))) return done(rc);
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
