// TV steepest descent (replaces src/tvmin/imgTVGradMin.cpp and
// imgTVGradMin3D.cpp): the gradient of the zero-padded isotropic TV functional
// (2-D 7-point / 3-D 13-point stencil), its FP64 norm, and the normalised step
// I -= alpha*d*v/||v|| with d = ||I - I0|| fixed for the call.
#include <mutex>

#include "reduce_device.hpp"
#include "tv.hpp"
#include "tv_stencil.hpp"

namespace asd {

namespace {

// s = alpha * d / ||v|| of the previous sub-iteration (reference: the step
// I -= alpha*d*v/||v||, imgTVGradMin.cpp:104-107).  ||v||^2 is folded from
// the previous kernel's block partials inside every block, in the fixed
// order of fold_partials, so each block gets the same value and no separate
// fold kernel runs between sub-iterations.
// nrm2_dev, when not NULL, holds ||v||^2 already (summed in index order by
// serial_sumsq_kernel) and the partials are not folded.
__device__ __forceinline__ TvStepVal tv_step_size(double alpha, const double *alpha_dev,
                                                  const double *d2, const double *partials_prev,
                                                  int nb_prev, const double *nrm2_dev)
{
#pragma clang fp contract(off)
    double nrm2[1];
    if (nrm2_dev)
        nrm2[0] = *nrm2_dev;
    else
        fold_rows<1>(partials_prev, nb_prev, nrm2);
    const double a = alpha_dev ? *alpha_dev : alpha;
    return tv_step_val(a * sqrt(*d2), sqrt(nrm2[0]));
}

// One TV gradient: v = grad TV(u) with FP64 block partials of ||v||^2.  With
// STEPPED it first applies the previous step to every value it reads and
// also stores the stepped image of its own voxel in u_next (the step fused
// into the next gradient; u_prev / v_prev are read-only here).
template <class T, class A, bool THREE_D, bool STEPPED>
__global__ __launch_bounds__(256) void tv_grad_kernel(const T *__restrict__ u,
                                                      const T *__restrict__ v_prev,
                                                      T *__restrict__ u_next, T *__restrict__ v,
                                                      int nx, int ny, int nz, double alpha,
                                                      const double *alpha_dev, const double *d2,
                                                      const double *partials_prev, int nb_prev,
                                                      double *partials, const int *skip,
                                                      const double *nrm2_dev)
{
    if (skip && *skip) return;
    TvSrc<T, A, STEPPED> U{u, v_prev, TvStepVal{0.0, 0.0, 1.0}, nx, ny, nz};
    const size_t n = (size_t)nx * ny * nz;
    double s = 0.0;
    // Each block walks one contiguous chunk of the image, and each XCD owns a
    // contiguous run of chunks (an eighth of the volume).  The chunks one
    // plane up and down belong to blocks of the same XCD that run at the same
    // time, so a stencil's z neighbours are read from that XCD's L2 while they
    // are still there instead of again from HBM.
    const size_t chunk = (n + (size_t)gridDim.x * 256 - 1) / ((size_t)gridDim.x * 256) * 256;
    const size_t i0 = (size_t)xcd_block(blockIdx.x, gridDim.x) * chunk;
    const size_t i1 = i0 + chunk < n ? i0 + chunk : n;
    VoxelCursor cur(i0 + threadIdx.x, nx, ny);
    size_t i = i0 + threadIdx.x;
    if constexpr (STEPPED) {
        // the first voxel's stencil is loaded before the previous norm is
        // folded, so the two memory round trips overlap
        constexpr int K = THREE_D ? 13 : 7;
        TvRaw<T, K> raw;
        if (i < i1) {
            const TvGather<T, A, K> gat{u, v_prev, i, cur.x, cur.y, cur.z, nx, ny, nz, raw, 0};
            (void)(THREE_D ? tv3_grad<A>(gat, cur.x, cur.y, cur.z)
                           : tv2_grad<A>(gat, cur.x, cur.y));
        }
        U.st = tv_step_size(alpha, alpha_dev, d2, partials_prev, nb_prev, nrm2_dev);
        // the gathered values are used only after the fold: the empty asm ties
        // them to the step size, which keeps the compiler from converting them
        // (and so waiting for their loads) before the fold starts
#pragma unroll
        for (int k = 0; k < K; k++)
            asm volatile("" : "+v"(raw.u[k]), "+v"(raw.v[k]) : "v"(U.st.s), "v"(U.st.nrm));
        if (i < i1) {
            const TvReplay<T, A, K> rep{raw, U.st, 0};
            A g = THREE_D ? tv3_grad<A>(rep, cur.x, cur.y, cur.z) : tv2_grad<A>(rep, cur.x, cur.y);
            T gt = T(g);
            v[i] = gt;
            u_next[i] = U.stored(i);
            s += (double)gt * (double)gt;
        }
        i += blockDim.x;
        cur.advance(256);
    }
    for (; i < i1; i += blockDim.x, cur.advance(256)) {
        const TvAt<T, A, STEPPED> at{U, i, cur.x, cur.y, cur.z};
        A g = THREE_D ? tv3_grad<A>(at, cur.x, cur.y, cur.z) : tv2_grad<A>(at, cur.x, cur.y);
        T gt = T(g);
        v[i] = gt;
        if constexpr (STEPPED) u_next[i] = U.stored(i);
        s += (double)gt * (double)gt;
    }
    double val[1] = {s};
    store_partials<1>(val, partials);
}

// The last step: u_out = u - s * v, then positivity when clamp.
template <class T>
__global__ __launch_bounds__(256) void tv_step_kernel(const T *__restrict__ u,
                                                      const T *__restrict__ v,
                                                      T *__restrict__ u_out, size_t n,
                                                      double alpha,
                                                      const double *__restrict__ alpha_dev,
                                                      const double *__restrict__ d2,
                                                      const double *__restrict__ partials_prev,
                                                      int nb_prev, int clamp, const int *skip,
                                                      const double *nrm2_dev)
{
#pragma clang fp contract(off)
    if (skip && *skip) return;
    const TvStepVal st = tv_step_size(alpha, alpha_dev, d2, partials_prev, nb_prev, nrm2_dev);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        if constexpr (sizeof(T) == sizeof(double)) {
            T w = st.apply<T>(u[i], v[i]);
            if (clamp && w < 0) w = 0;
            u_out[i] = w;
        } else {
            // clamped in FP64 before the rounding to T
            double w = (double)u[i] - st.s * (double)v[i];
            if (clamp && w < 0) w = 0;
            u_out[i] = T(w);
        }
    }
}

template <class T>
__global__ __launch_bounds__(256) void sqdiff_kernel(const T *__restrict__ a,
                                                     const T *__restrict__ b, size_t n,
                                                     double *partials)
{
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        double e = (double)a[i] - (b ? (double)b[i] : 0.0);
        s += e * e;
    }
    double val[1] = {s};
    store_partials<1>(val, partials);
}

unsigned stream_blocks(size_t n)
{
    size_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    return (unsigned)(b ? b : 1);
}

}  // namespace

template <class T>
int launch_sqdiff(const T *a, const T *b, size_t n, TvScratch &s, int slot, hipStream_t st)
{
    unsigned blocks = red_blocks(n);
    sqdiff_kernel<T><<<blocks, RED_THREADS, 0, st>>>(a, b, n, s.partials);
    ASD_LAUNCH_CHECK("sqdiff_kernel");
    fold_partials<1><<<1, RED_THREADS, 0, st>>>(s.partials, (int)blocks, s.scalars + slot, nullptr);
    ASD_LAUNCH_CHECK("fold_partials");
    return ASD_OK;
}

template <class T>
int launch_serial_sumsq(const T *a, const T *b, size_t n, double *out, hipStream_t st,
                        const int *skip)
{
    serial_sumsq_kernel<T><<<1, 64, 0, st>>>(a, b, n, out, skip);
    ASD_LAUNCH_CHECK("serial_sumsq_kernel");
    return ASD_OK;
}

template <class T, class A>
int tv_descent(T *u, const T *img_old, T *v, T *u2, T *v2, int nx, int ny, int nz, bool three_d,
               double alpha, const double *alpha_dev, int nTVIter, int pos, TvScratch &s,
               hipStream_t st, bool have_d2, const int *skip, bool serial_norms)
{
    const size_t n = (size_t)nx * ny * nz;
    int rc;
    if (!have_d2) {
        rc = serial_norms ? launch_serial_sumsq<T>(u, img_old, n, s.scalars + TV_SLOT_D2, st, skip)
                          : launch_sqdiff<T>(u, img_old, n, s, TV_SLOT_D2, st);
        if (rc) return rc;
    }
    if (nTVIter <= 0) return ASD_OK;
    const unsigned gb = red_blocks(n), ub = stream_blocks(n);
    double *d2 = s.scalars + TV_SLOT_D2;
    // serial norms: ||v||^2 of each gradient summed in index order into one
    // slot, read by the next kernel instead of the block partials
    double *nrm2 = serial_norms ? s.scalars + TV_SLOT_NRM2 : nullptr;
    // gradient partials alternate between two rows: kernel i writes one while
    // reading kernel i-1's in the other
    double *pp[2] = {s.partials, s.partials + RED_MAX_BLOCKS};
    // (u, v) holds sub-iteration i's image and gradient, alternating with (u2, v2)
    T *cu = u, *cv = v, *nu = u2, *nv = v2;
    for (int it = 0; it < nTVIter; it++) {
        double *out = pp[it & 1];
        const double *prev = pp[(it + 1) & 1];
        if (it == 0) {
            if (three_d)
                tv_grad_kernel<T, A, true, false><<<gb, RED_THREADS, 0, st>>>(
                    cu, nullptr, nullptr, cv, nx, ny, nz, alpha, alpha_dev, d2, nullptr, 0, out,
                    skip, nullptr);
            else
                tv_grad_kernel<T, A, false, false><<<gb, RED_THREADS, 0, st>>>(
                    cu, nullptr, nullptr, cv, nx, ny, 1, alpha, alpha_dev, d2, nullptr, 0, out,
                    skip, nullptr);
        } else {
            // step of sub-iteration it-1 fused into gradient it
            if (three_d)
                tv_grad_kernel<T, A, true, true><<<gb, RED_THREADS, 0, st>>>(
                    cu, cv, nu, nv, nx, ny, nz, alpha, alpha_dev, d2, prev, (int)gb, out, skip,
                    nrm2);
            else
                tv_grad_kernel<T, A, false, true><<<gb, RED_THREADS, 0, st>>>(
                    cu, cv, nu, nv, nx, ny, 1, alpha, alpha_dev, d2, prev, (int)gb, out, skip,
                    nrm2);
            T *t1 = cu, *t2 = cv;
            cu = nu;
            cv = nv;
            nu = t1;
            nv = t2;
        }
        ASD_LAUNCH_CHECK("tv_grad_kernel");
        if (serial_norms && (rc = launch_serial_sumsq<T>(cv, nullptr, n, nrm2, st, skip)))
            return rc;
    }
    // last step (with positivity) lands in the caller's buffer u
    tv_step_kernel<T><<<ub, 256, 0, st>>>(cu, cv, u, n, alpha, alpha_dev, d2,
                                          pp[(nTVIter - 1) & 1], (int)gb, pos == 1 ? 1 : 0,
                                          skip, nrm2);
    ASD_LAUNCH_CHECK("tv_step_kernel");
    return ASD_OK;
}

template int launch_sqdiff<float>(const float *, const float *, size_t, TvScratch &, int,
                                  hipStream_t);
template int launch_sqdiff<double>(const double *, const double *, size_t, TvScratch &, int,
                                   hipStream_t);
template int launch_serial_sumsq<float>(const float *, const float *, size_t, double *,
                                        hipStream_t, const int *);
template int launch_serial_sumsq<double>(const double *, const double *, size_t, double *,
                                         hipStream_t, const int *);
template int tv_descent<float, float>(float *, const float *, float *, float *, float *, int, int,
                                      int, bool, double, const double *, int, int, TvScratch &,
                                      hipStream_t, bool, const int *, bool);
template int tv_descent<double, double>(double *, const double *, double *, double *, double *,
                                        int, int, int, bool, double, const double *, int, int,
                                        TvScratch &, hipStream_t, bool, const int *, bool);

int TvScratch::init()
{
    int rc;
    if ((rc = buf.ensure(sizeof(double) * (RED_MAX_SUMS * RED_MAX_BLOCKS + TV_NSLOTS))))
        return rc;
    partials = buf.as<double>();
    scalars = partials + RED_MAX_SUMS * RED_MAX_BLOCKS;
    ASD_HIP(hipMemset(buf.p, 0, buf.bytes));
    return ASD_OK;
}

// ---------------------------------------------------------------------------
// Host entry points (the imgTVGradMin / imgTVGradMin3D compute).
// ---------------------------------------------------------------------------
namespace {

struct TvCtx {
    std::mutex mu;
    hipStream_t stream = nullptr;
    DeviceBuf img, old, v, u2, v2;
    TvScratch scratch;
    bool ready = false;
    int init()
    {
        if (ready) return ASD_OK;
        if (!stream) ASD_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        int rc = scratch.init();
        if (rc) return rc;
        ready = true;
        return ASD_OK;
    }
};

TvCtx &tvctx()
{
    static TvCtx *c = new TvCtx();
    return *c;
}

int tv_host(const double *img, int nx, int ny, int nz, bool three_d, double alpha,
            int nTVIter, int pos, const double *img_old, double *out)
{
    clear_error();
    ASD_ARG(img && out, "NULL image pointer");
    ASD_ARG(nx > 0 && ny > 0 && nz > 0, "image dimensions must be positive");
    TvCtx &c = tvctx();
    std::lock_guard<std::mutex> lk(c.mu);
    int rc = c.init();
    if (rc) return rc;
    size_t n = (size_t)nx * ny * nz;
    if ((rc = c.img.ensure(n * 8)) || (rc = c.v.ensure(n * 8)) || (rc = c.u2.ensure(n * 8)) ||
        (rc = c.v2.ensure(n * 8)))
        return rc;
    if (img_old && (rc = c.old.ensure(n * 8))) return rc;
    ASD_HIP(hipMemcpyAsync(c.img.p, img, n * 8, hipMemcpyHostToDevice, c.stream));
    if (img_old)
        ASD_HIP(hipMemcpyAsync(c.old.p, img_old, n * 8, hipMemcpyHostToDevice, c.stream));
    // the drop-in forms sum d^2 and ||v||^2 in the reference's loop order
    // (imgTVGradMin.cpp:58-59, :96; imgTVGradMin3D.cpp:150-153), so the
    // result rounds like the reference's
    rc = tv_descent<double, double>(c.img.as<double>(), img_old ? c.old.as<double>() : nullptr,
                                    c.v.as<double>(), c.u2.as<double>(), c.v2.as<double>(), nx, ny,
                                    nz, three_d, alpha, nullptr, nTVIter, pos, c.scratch, c.stream,
                                    false, nullptr, true);
    if (rc) return rc;
    ASD_HIP(hipMemcpyAsync(out, c.img.p, n * 8, hipMemcpyDeviceToHost, c.stream));
    ASD_HIP(hipStreamSynchronize(c.stream));
    return ASD_OK;
}

}  // namespace
}  // namespace asd

using namespace asd;

extern "C" {

int asd_tv_min_2d(const double *img, int rowN, int colN, double alpha, int nTVIter, int pos,
                  const double *img_old, double *out)
{
    return tv_host(img, rowN, colN, 1, false, alpha, nTVIter, pos, img_old, out);
}

int asd_tv_min_3d(const double *img, int Nx, int Ny, int Nz, double alpha, int nTVIter,
                  int pos, const double *img_old, double *out)
{
    return tv_host(img, Nx, Ny, Nz, true, alpha, nTVIter, pos, img_old, out);
}

int asd_tv_min_device(double *img, const double *img_old, int nx, int ny, int nz, int three_d,
                      double alpha, int nTVIter, int pos, void *stream)
{
    clear_error();
    ASD_ARG(img, "NULL image pointer");
    ASD_ARG(nx > 0 && ny > 0 && nz > 0, "image dimensions must be positive");
    static thread_local TvScratch scratch;
    static thread_local DeviceBuf v, u2, v2;
    int rc;
    if (!scratch.partials && (rc = scratch.init())) return rc;
    size_t n = (size_t)nx * ny * nz;
    if ((rc = v.ensure(n * 8)) || (rc = u2.ensure(n * 8)) || (rc = v2.ensure(n * 8))) return rc;
    return tv_descent<double, double>(img, img_old, v.as<double>(), u2.as<double>(),
                                      v2.as<double>(), nx, ny, three_d ? nz : 1, three_d != 0,
                                      alpha, nullptr, nTVIter, pos, scratch, (hipStream_t)stream,
                                      false);
}

}  // extern "C"
