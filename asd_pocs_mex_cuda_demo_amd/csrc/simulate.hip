// Synthetic reconstruction inputs made on the device (the step before the
// reconstruction: the reference builds them in MATLAB, demo_test.m:35-52):
//
//   phantom_kernel    ellipsoid phantom rasterised on the voxel grid, the same
//                     FP64 expressions in the same order as the package's CPU
//                     phantom (phantoms.py), so both give the same volume
//   poisson_log       counts ~ Poisson(I0 * exp(-p)), p' = -ln(max(counts,1)/I0)
//                     with a counter-based generator (Philox4x32-10) and, for
//                     means >= 10, Hormann's transformed rejection (PTRS)
//   asd_plan_simulate the phantom into the plan's RMSE reference, every view
//                     forward projected into an FP64 staging sinogram, noise,
//                     then encoded into the plan's sinogram storage exactly as
//                     asd_plan_set_projections encodes a host sinogram
#include <hip/hip_fp16.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "plan.hpp"
#include "projector_kernels.hpp"
#include "reduce_device.hpp"

namespace asd {

namespace {

struct Ellipsoid {
    double A, a, b, c, x0, y0, z0;
    double cos_t, sin_t;     // 2-D: rotation by phi about the ellipse's centre
    double m[9];             // 3-D: phantom3d.m's Euler matrix of (phi, theta, psi)
    double asq, bsq, csq;    // a^2, b^2, c^2
};

// One voxel per thread: the sum of A over the ellipsoids holding the voxel
// centre, in table order, times scale, rounded to FP32, with the FP64
// expressions of the CPU rasteriser (phantoms.py) in the same order.
//  - 2-D (nz == 1): x = r(ix), y = r(iy) with r(i) = (i - (n-1)/2) / ((n-1)/2),
//    each ellipse rotated by phi about its centre.
//  - 3-D: phantom3d.m:76-108 - the point (x, y, z) = (r(iy), r(ix), r(iz +
//    (nx - nz) floor/ 2)) (MATLAB's meshgrid puts x along the second index)
//    rotated by the Euler matrix, then (p - offset)^2 / axis^2 summed.
__global__ __launch_bounds__(256) void phantom_kernel(const Ellipsoid *__restrict__ E, int ne,
                                                      int nx, int ny, int nz, double scale,
                                                      float *__restrict__ out)
{
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t N = (size_t)nx * ny * nz;
    if (i >= N) return;
    const int ix = (int)(i % nx), iy = (int)((i / nx) % ny), iz = (int)(i / ((size_t)nx * ny));
    const double hx = (nx - 1) / 2.0, hy = (ny - 1) / 2.0;
    double v = 0.0;
    if (nz == 1) {
        const double X = ((double)ix - hx) / hx, Y = ((double)iy - hy) / hy;
        for (int k = 0; k < ne; k++) {
            const Ellipsoid e = E[k];
            const double dx = X - e.x0, dy = Y - e.y0;
            const double xr = dx * e.cos_t + dy * e.sin_t;
            const double yr = (-dx) * e.sin_t + dy * e.cos_t;
            const double tx = xr / e.a, ty = yr / e.b;
            if (tx * tx + ty * ty <= 1.0) v += e.A;
        }
    } else {
        const int off = nx >= nz ? (nx - nz) / 2 : -((nz - nx + 1) / 2);
        const double x = ((double)iy - hy) / hy, y = ((double)ix - hx) / hx;
        const double z = ((double)(iz + off) - hx) / hx;
        for (int k = 0; k < ne; k++) {
            const Ellipsoid &e = E[k];
            const double p1 = e.m[0] * x + e.m[1] * y + e.m[2] * z;
            const double p2 = e.m[3] * x + e.m[4] * y + e.m[5] * z;
            const double p3 = e.m[6] * x + e.m[7] * y + e.m[8] * z;
            const double d1 = p1 - e.x0, d2 = p2 - e.y0, d3 = p3 - e.z0;
            const double q = (d1 * d1) / e.asq + (d2 * d2) / e.bsq + (d3 * d3) / e.csq;
            if (q <= 1.0) v += e.A;
        }
    }
    out[i] = (float)(v * scale);
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011): ten rounds of two 32x32->64 multiplications
// with a Weyl-sequence key schedule.  Counter = (element, draw), key = seed.
struct Philox {
    uint32_t k0, k1;
    __device__ __forceinline__ void block(uint32_t (&c)[4]) const
    {
        uint32_t a = k0, b = k1;
#pragma unroll
        for (int r = 0; r < 10; r++) {
            const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
            const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
            const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ a;
            const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ b;
            c[0] = n0;
            c[1] = (uint32_t)p1;
            c[2] = n2;
            c[3] = (uint32_t)p0;
            a += 0x9E3779B9u;
            b += 0xBB67AE85u;
        }
    }
};

// Uniform doubles in (0, 1) from one element's stream, two per Philox block.
struct Stream {
    Philox ph;
    uint64_t elem;
    uint32_t draw = 0;
    uint32_t buf[4];
    int left = 0;
    __device__ __forceinline__ double uniform()
    {
        if (left == 0) {
            buf[0] = (uint32_t)elem;
            buf[1] = (uint32_t)(elem >> 32);
            buf[2] = draw++;
            buf[3] = 0x5A17u;
            ph.block(buf);
            left = 2;
        }
        const int j = 2 - left;
        left--;
        const uint64_t bits = ((uint64_t)buf[2 * j] << 32) | buf[2 * j + 1];
        return ((double)(bits >> 11) + 0.5) * 0x1.0p-53;
    }
};

// Poisson(lam) deviate.  lam < 10: the product of uniforms (Knuth).  lam >=
// 10: PTRS, the transformed rejection with squeeze of W. Hormann, "The
// transformed rejection method for generating Poisson random variables",
// Insurance: Mathematics and Economics 12 (1993).  Every loop is bounded; the
// bounds are far beyond what the acceptance rates ever need (a bound reached
// returns the rounded mean).
__device__ double poisson(double lam, Stream &s)
{
    if (!(lam > 0.0)) return 0.0;
    if (lam < 10.0) {
        const double L = exp(-lam);
        double p = 1.0;
        for (int k = 0; k < 1000; k++) {
            p *= s.uniform();
            if (p <= L) return (double)k;
        }
        return floor(lam + 0.5);
    }
    const double slam = sqrt(lam), loglam = log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (int it = 0; it < 1000; it++) {
        const double U = s.uniform() - 0.5, V = s.uniform();
        const double us = 0.5 - fabs(U);
        const double k = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1.0))
            return k;
    }
    return floor(lam + 0.5);
}

// p' = -ln(max(counts, 1) / I0), counts ~ Poisson(I0 * exp(-p)), in place
// over n FP64 values starting at global element `first`.
__global__ __launch_bounds__(256) void poisson_log_kernel(double *__restrict__ P, size_t n,
                                                          size_t first, double I0, uint32_t k0,
                                                          uint32_t k1)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    Stream s{Philox{k0, k1}, first + i};
    const double c = poisson(I0 * exp(-P[i]), s);
    P[i] = -log((c > 1.0 ? c : 1.0) / I0);
}

__global__ __launch_bounds__(256) void f32_to_f64(const float *__restrict__ a,
                                                  double *__restrict__ b, size_t n)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}

// Per-view minimum and maximum of an (M x nv) FP64 sinogram: one block per
// view writes (lo, hi) to mm[2 v], mm[2 v + 1].
__global__ __launch_bounds__(256) void view_minmax(const double *__restrict__ P, size_t M,
                                                   double *__restrict__ mm)
{
    const double *v = P + (size_t)blockIdx.x * M;
    double lo = v[0], hi = v[0];
    for (size_t i = threadIdx.x; i < M; i += 256) {
        lo = fmin(lo, v[i]);
        hi = fmax(hi, v[i]);
    }
    __shared__ double sl[256], sh[256];
    sl[threadIdx.x] = lo;
    sh[threadIdx.x] = hi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sl[threadIdx.x] = fmin(sl[threadIdx.x], sl[threadIdx.x + s]);
            sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        mm[2 * blockIdx.x] = sl[0];
        mm[2 * blockIdx.x + 1] = sh[0];
    }
}

// The codecs' encoders (the same arithmetic as asd_plan_set_projections).
__global__ __launch_bounds__(256) void encode_f32(const double *__restrict__ P,
                                                  float *__restrict__ out, size_t n)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = (float)P[i];
}

__global__ __launch_bounds__(256) void encode_f16(const double *__restrict__ P,
                                                  __half *__restrict__ out, size_t n)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = __float2half_rn((float)P[i]);
}

__global__ __launch_bounds__(256) void encode_u16(const double *__restrict__ P,
                                                  uint16_t *__restrict__ out, size_t M, size_t n,
                                                  const double *__restrict__ lo_step)
{
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const size_t v = i / M;
    const double lo = lo_step[2 * v], step = lo_step[2 * v + 1];
    const double t = step > 0 ? (P[i] - lo) / step : 0.0;
    const long r = lround(t);
    out[i] = (uint16_t)(r < 0 ? 0 : r > 65535 ? 65535 : r);
}

unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

}  // namespace asd

using namespace asd;

extern "C" {

int asd_plan_simulate(asd_plan *p, const double *ellipsoids, int n_ell, double scale,
                      double I0_photons, unsigned long long seed)
{
    clear_error();
    ASD_ARG(p, "NULL plan");
    ASD_ARG(n_ell >= 0 && (n_ell == 0 || ellipsoids), "bad ellipsoid table");
    ASD_ARG(I0_photons >= 0.0, "I0_photons must be >= 0 (0: no noise)");
    const asd_para &pa = p->para;
    ASD_ARG(pa.nx > 1 && pa.ny > 1, "the phantom grid needs nx, ny >= 2");
    int rc;
    // ellipsoid table: A, a, b, c, x0, y0, z0, phi, theta, psi (degrees) per
    // row (phantom3d.m's layout); the rotations are formed here with the
    // expressions of phantoms.py (2-D: phi * (pi / 180); 3-D: the Euler
    // matrix of phantom3d.m:94-101, each angle as deg * pi / 180)
    std::vector<Ellipsoid> E(n_ell);
    for (int k = 0; k < n_ell; k++) {
        const double *r = ellipsoids + 10 * k;
        Ellipsoid &e = E[k];
        e.A = r[0]; e.a = r[1]; e.b = r[2]; e.c = r[3];
        e.x0 = r[4]; e.y0 = r[5]; e.z0 = r[6];
        const double t = r[7] * (M_PI / 180.0);
        e.cos_t = std::cos(t);
        e.sin_t = std::sin(t);
        const double ph = r[7] * M_PI / 180, th = r[8] * M_PI / 180, ps = r[9] * M_PI / 180;
        const double cph = std::cos(ph), sph = std::sin(ph), cth = std::cos(th),
                     sth = std::sin(th), cps = std::cos(ps), sps = std::sin(ps);
        const double m[9] = {cps * cph - cth * sph * sps, cps * sph + cth * cph * sps, sps * sth,
                             -sps * cph - cth * sph * cps, -sps * sph + cth * cph * cps,
                             cps * sth, sth * sph, -sth * cph, cth};
        for (int j = 0; j < 9; j++) e.m[j] = m[j];
        e.asq = r[1] * r[1];
        e.bsq = r[2] * r[2];
        e.csq = r[3] * r[3];
    }
    DeviceBuf dE, stage, mm, srb, img32;
    // the phantom in FP32 (the reference-arithmetic plan keeps its RMSE
    // reference in FP64 and rasterises into a separate FP32 image first)
    DeviceBuf &img = p->ref64() ? img32 : p->I0;
    if ((rc = dE.ensure(sizeof(Ellipsoid) * (n_ell > 0 ? n_ell : 1))) ||
        (rc = img.ensure(p->N * 4)) || (rc = p->I0.ensure(p->N * p->img_bytes())) ||
        (rc = srb.ensure(p->M * 4)))
        return rc;
    if (n_ell) ASD_HIP(hipMemcpyAsync(dE.p, E.data(), sizeof(Ellipsoid) * n_ell,
                                      hipMemcpyHostToDevice, p->stream));
    phantom_kernel<<<blocks_for(p->N), 256, 0, p->stream>>>(dE.as<Ellipsoid>(), n_ell, pa.nx,
                                                            pa.ny, pa.nz, scale,
                                                            img.as<float>());
    ASD_LAUNCH_CHECK("phantom_kernel");
    if (p->ref64()) {
        f32_to_f64<<<blocks_for(p->N), 256, 0, p->stream>>>(img.as<float>(), p->I0.as<double>(),
                                                            p->N);
        ASD_LAUNCH_CHECK("f32_to_f64");
    }
    // every view forward projected (the drop-in projector's kernel and lane
    // count) into an FP64 staging sinogram, then noise
    const size_t M = p->M, n = (size_t)p->nv * M;
    if ((rc = stage.ensure(n * 8)) || (rc = mm.ensure(sizeof(double) * 2 * p->nv))) return rc;
    for (int iv = 0; iv < p->nv; iv++) {
        const ViewGeom &g = p->views[iv];
        const float *yd = p->ydet.as<float>() + (size_t)iv * pa.na;
        const float *zd = p->zdet.as<float>() + (size_t)iv * pa.nb;
        if ((rc = launch_ax(g, img.as<float>(), p->cor.as<float>(), srb.as<float>(), yd, zd,
                            p->stream)))
            return rc;
        f32_to_f64<<<blocks_for(M), 256, 0, p->stream>>>(p->cor.as<float>(),
                                                          stage.as<double>() + iv * M, M);
        ASD_LAUNCH_CHECK("f32_to_f64");
    }
    if (I0_photons > 0.0) {
        poisson_log_kernel<<<blocks_for(n), 256, 0, p->stream>>>(
            stage.as<double>(), n, 0, I0_photons, (uint32_t)seed, (uint32_t)(seed >> 32));
        ASD_LAUNCH_CHECK("poisson_log_kernel");
    }
    // encode into the plan's storage
    if (pa.precision == 1 || pa.precision == 3) {
        view_minmax<<<p->nv, 256, 0, p->stream>>>(stage.as<double>(), M, mm.as<double>());
        ASD_LAUNCH_CHECK("view_minmax");
        std::vector<double> h(2 * p->nv);
        ASD_HIP(hipMemcpyAsync(h.data(), mm.p, h.size() * 8, hipMemcpyDeviceToHost, p->stream));
        ASD_HIP(hipStreamSynchronize(p->stream));
        p->u16_lo.assign(p->nv, 0.0);
        p->u16_step.assign(p->nv, 0.0);
        for (int iv = 0; iv < p->nv; iv++) {
            const double lo = h[2 * iv], hi = h[2 * iv + 1];
            const double step = hi > lo ? (hi - lo) / 65535.0 : 0.0;
            p->u16_lo[iv] = h[2 * iv] = lo;
            p->u16_step[iv] = h[2 * iv + 1] = step;
        }
        ASD_HIP(hipMemcpyAsync(mm.p, h.data(), h.size() * 8, hipMemcpyHostToDevice, p->stream));
        encode_u16<<<blocks_for(n), 256, 0, p->stream>>>(stage.as<double>(),
                                                          p->P.as<uint16_t>(), M, n,
                                                          mm.as<double>());
        ASD_LAUNCH_CHECK("encode_u16");
        plan_drop_graph(p);  // the view scales are kernel arguments
    } else if (pa.precision == 2 || pa.precision == 5) {
        encode_f16<<<blocks_for(n), 256, 0, p->stream>>>(stage.as<double>(), p->P.as<__half>(),
                                                          n);
        ASD_LAUNCH_CHECK("encode_f16");
    } else if (pa.precision == 4) {
        ASD_HIP(hipMemcpyAsync(p->P.p, stage.p, n * 8, hipMemcpyDeviceToDevice, p->stream));
    } else {
        encode_f32<<<blocks_for(n), 256, 0, p->stream>>>(stage.as<double>(), p->P.as<float>(),
                                                          n);
        ASD_LAUNCH_CHECK("encode_f32");
    }
    ASD_HIP(hipStreamSynchronize(p->stream));  // the staging buffers are freed on return
    p->have_P = true;
    if (!p->have_I0) plan_drop_graph(p);  // the graph bakes in the I0 pointer
    p->have_I0 = true;
    p->I0_sim = true;
    return ASD_OK;
}

int asd_plan_get_projections(asd_plan *p, double *P)
{
    clear_error();
    ASD_ARG(p && P, "NULL argument");
    ASD_ARG(p->have_P, "the plan has no projections");
    const size_t n = (size_t)p->nv * p->M;
    if (p->para.precision == 0) {
        std::vector<float> h(n);
        ASD_HIP(hipMemcpy(h.data(), p->P.p, n * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) P[i] = h[i];
    } else if (p->para.precision == 2 || p->para.precision == 5) {
        std::vector<__half> h(n);
        ASD_HIP(hipMemcpy(h.data(), p->P.p, n * 2, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) P[i] = (double)__half2float(h[i]);
    } else if (p->para.precision == 4) {
        ASD_HIP(hipMemcpy(P, p->P.p, n * 8, hipMemcpyDeviceToHost));
    } else {
        std::vector<uint16_t> h(n);
        ASD_HIP(hipMemcpy(h.data(), p->P.p, n * 2, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) {
            const size_t v = i / p->M;
            P[i] = p->u16_lo[v] + p->u16_step[v] * (double)h[i];
        }
    }
    return ASD_OK;
}

int asd_plan_get_reference(asd_plan *p, double *I0)
{
    clear_error();
    ASD_ARG(p && I0, "NULL argument");
    ASD_ARG(p->have_I0, "the plan has no RMSE reference image");
    if (p->ref64()) {
        ASD_HIP(hipMemcpy(I0, p->I0.p, p->N * 8, hipMemcpyDeviceToHost));
        return ASD_OK;
    }
    std::vector<float> h(p->N);
    ASD_HIP(hipMemcpy(h.data(), p->I0.p, p->N * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < p->N; i++) I0[i] = h[i];
    return ASD_OK;
}

}  // extern "C"
