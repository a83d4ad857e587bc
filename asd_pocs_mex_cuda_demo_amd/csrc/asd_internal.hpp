// Internal declarations of libasdpocs: error plumbing, geometry records and
// the device-side geometry helpers shared by the projector and driver kernels.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "asdpocs.h"

namespace asd {

// ---------------------------------------------------------------------------
// Errors: every HIP call is checked; failures become a status code plus a
// thread-local message readable through asd_last_error().
// ---------------------------------------------------------------------------
void set_error(int code, const std::string &msg);
void clear_error();
int last_status();

#define ASD_HIP(call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            ::asd::set_error(ASD_ERR_HIP, std::string(#call) + ": " +              \
                                              hipGetErrorString(e_));              \
            return ASD_ERR_HIP;                                                    \
        }                                                                          \
    } while (0)

#define ASD_LAUNCH_CHECK(what)                                                     \
    do {                                                                           \
        hipError_t e_ = hipGetLastError();                                         \
        if (e_ != hipSuccess) {                                                    \
            ::asd::set_error(ASD_ERR_HIP, std::string(what) + ": " +               \
                                              hipGetErrorString(e_));              \
            return ASD_ERR_HIP;                                                    \
        }                                                                          \
    } while (0)

#define ASD_ARG(cond, msg)                                                         \
    do {                                                                           \
        if (!(cond)) {                                                             \
            ::asd::set_error(ASD_ERR_ARG, msg);                                    \
            return ASD_ERR_ARG;                                                    \
        }                                                                          \
    } while (0)

// Unsigned division by a runtime constant d >= 1 as a multiply-high, an add
// and shifts (T. Granlund and P. Montgomery, "Division by invariant integers
// using multiplication", PLDI 1994, Fig. 4.1): exact for every 32-bit
// dividend, and a handful of instructions instead of a general division.
struct FastDiv {
    unsigned m;
    int s1, s2;
};

__host__ __device__ inline FastDiv make_fastdiv(unsigned d)
{
    if (d == 0) d = 1;  // sizes are validated separately; never divide by zero here
    int l = 0;          // ceil(log2 d)
    while ((1ull << l) < d) l++;
    FastDiv f;
    f.m = (unsigned)((((1ull << 32) * ((1ull << l) - d)) / d) + 1);
    f.s1 = l < 1 ? l : 1;
    f.s2 = l > 1 ? l - 1 : 0;
    return f;
}

__host__ __device__ __forceinline__ unsigned fastdiv(unsigned n, const FastDiv &f)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const unsigned t = __umulhi(f.m, n);
#else
    const unsigned t = (unsigned)(((unsigned long long)f.m * n) >> 32);
#endif
    return (t + ((n - t) >> f.s1)) >> f.s2;
}

// ---------------------------------------------------------------------------
// Geometry of one view in voxel units (what the reference gateways compute,
// src/AxAtx/Ax_main_test.cpp:138-161).
// ---------------------------------------------------------------------------
struct ViewGeom {
    float cos_phi, sin_phi;  // rotation of this view
    float SO, OD;            // source-isocentre, isocentre-detector (voxels)
    float scale;             // voxel size in mm (outputs are multiplied by it)
    float dz;                // axial voxel pitch (1 in the reference)
    float dy_det, dz_det;    // detector pitch (voxels), used by the backprojector footprint
    int nx, ny, nz, na, nb;
    float fp_d;              // footprint half-extent sqrt((1 + dz^2) / 2) (geom_finish)
    float fp_cx[4], fp_cy[4];  // voxel corner offsets in the view frame (bounds only)
    float fp_rdy, fp_rdz;    // 1 / dy_det, 1 / dz_det (bounds only)
    int tight;               // footprints may be cut to the voxel's shadow (view_tight)
    int general;             // every ray is a general ray with exact face offsets (view_general)
    FastDiv div_nx, div_ny;  // voxel index -> (x, row) and row -> (y, z) (geom_finish)
    FastDiv div_na;          // bin index -> (ia, ib) (geom_finish)
    // z-window [z0, z0 + nzs) of the volume the buffers hold (geom_finish:
    // the whole volume).  The slab entry points narrow it: the forward
    // kernels then sum only the pieces in those slices (a slab's partial ray
    // sums, read from a slab-resident volume) and atx_kernel evaluates only
    // those slices' voxels; geometry stays the whole volume's (nz).
    int z0, nzs;
};

// Per-ray record for the backprojector: the six slope quotients the
// reference's find_l_3d forms per (voxel, ray) (cu:142-196) depend only on the
// ray, so they are formed once per ray with the same IEEE divisions.
struct alignas(32) RayRec {
    float s_ba, s_ab, s_ca, s_cb, s_ac, s_bc;
    int mode;   // RAY_* below
    // General rays with exact face offsets (RAY_X_FIRST, RAY_Y_FIRST): bytes
    // 0, 1 and 3 hold, for x, y and z, the top byte of the FP32 coordinate
    // -sign(d_k) / 2 of the face of that axis the ray enters through, with
    // (a, b, c) = d the ray's direction (0xbf: -1/2, 0x3f: +1/2).  Zero for
    // the other modes.
    unsigned side;
};

// find_l_3d dispatch (cu:131-200 and find_l cu:26-113)
enum : int {
    RAY_X_FIRST = 0,      // a,b != 0, |a| >  |b|: x faces, y faces, z faces
    RAY_Y_FIRST = 1,      // a,b != 0, |a| <= |b|: y faces, x faces, z faces
    RAY_A0_B0 = 2,        // a == 0, b == 0: 2-D (y,z) chord, axis-parallel in y -> dz
    RAY_A0_C0 = 3,        // a == 0, c == 0: |pz| <= dz/2 -> dy
    RAY_A0_YZ_U = 4,      // a == 0: (y,z) chord, |b| >  |c|
    RAY_A0_YZ_W = 5,      // a == 0: (y,z) chord, |b| <= |c|
    RAY_B0_C0 = 6,        // b == 0, c == 0: |pz| <= dz/2 -> dx
    RAY_B0_XZ_U = 7,      // b == 0: (x,z) chord, |a| >  |c|
    RAY_B0_XZ_W = 8,      // b == 0: (x,z) chord, |a| <= |c|
    // RAY_X_FIRST / RAY_Y_FIRST for a ray where some slope s has s * h
    // inexact (h = 0.5 and dz / 2; a subnormal product, or dz != 1): the
    // face end points c -+ s*h are formed from the rounded products
    RAY_X_FIRST_P = 9,
    RAY_Y_FIRST_P = 10,
    // RAY_X_FIRST / RAY_Y_FIRST for a ray in the z = 0 face of a one-slice
    // volume (c == 0, nz == 1, dz == 1): the in-plane chord (chord2)
    RAY_FLAT_X = 11,
    RAY_FLAT_Y = 12,
};

__host__ __device__ __forceinline__ bool ray_general(int mode)
{
    return mode <= RAY_Y_FIRST || mode >= RAY_X_FIRST_P;
}

// Host helpers shared by the translation units.
struct DeviceBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t n);   // grow-only
    void release();
    ~DeviceBuf() { release(); }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

int launch_ray_table(const ViewGeom &g, const float *y_det, const float *z_det, RayRec *rt,
                     hipStream_t s);
int launch_ax(const ViewGeom &g, const float *x, float *y, float *sr, const float *y_det,
              const float *z_det, hipStream_t s);
int launch_atx(const ViewGeom &g, float *x, const float *y, float *sc, const float *y_det,
               const float *z_det, const RayRec *rt, hipStream_t s);

// Correctly rounded square root of a finite x >= 2^-96, from the hardware's
// reciprocal square root y ~ 1/sqrt(x) by one residual step: s0 = RN(x*y),
// r = x - s0*s0 (exact in an FMA), s = RN(s0 + r*(y/2)).  This is
// Markstein's final step for the square root (P. Markstein, IBM J. Res.
// Dev. 34(1), 1990; Muller et al., Handbook of Floating-Point Arithmetic,
// the FMA-based square-root algorithms).  On gfx950 its result equals the
// correctly rounded sqrtf for every float in [2^-96, FLT_MAX]: a GPU test
// compares the two on each of those 1.88e9 inputs (tests/gpu_checks).  It
// costs one quarter-rate rsq and four full-rate instructions, against the
// quarter-rate sqrt, two neighbours, two residuals and two selects of the
// compiler's correctly rounded expansion.
__host__ __device__ __forceinline__ float sqrt_finite_ge_tiny(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float y = __builtin_amdgcn_rsqf(x);
    const float s0 = x * y;
    const float h = 0.5f * y;
    const float r = __builtin_fmaf(-s0, s0, x);
    return __builtin_fmaf(r, h, s0);
#else
    return sqrtf(x);
#endif
}

// Correctly rounded square root of x >= 2^-96, +-0, +inf or NaN:
// sqrt_finite_ge_tiny, with +-0 and +inf returned as their own roots (class
// mask: -0, +0, +inf), for which its residual step would give NaN.
__host__ __device__ __forceinline__ float sqrt_ge_tiny(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float t = sqrt_finite_ge_tiny(x);
    return __builtin_amdgcn_classf(x, 0x020 | 0x040 | 0x200) ? x : t;
#else
    return sqrtf(x);
#endif
}

// Lanes cooperating on one ray of the forward projector.
int ax_lanes_per_ray(int na, int nb, int nu);

// Launch shape of the forward projector for one view: lanes per ray (per
// bundle when rows > 1) and detector rows per bundle (1: ax_kernel, one ray
// per lane group; 2 or 3: ax_rows_kernel at 8 lanes).  The shape changes only the order in
// which a ray's slab products are summed (and so the last bits of the sum),
// never which products there are.
struct AxShape {
    int lpr = 1, rows = 1;
};
AxShape ax_shape(int na, int nb, int nu);

// XCD-aware block order: hardware hands consecutive workgroups to the 8 XCDs
// round-robin; remap so that each XCD walks a contiguous range of tiles and
// neighbouring tiles share its L2.
__host__ __device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nblocks)
{
    constexpr unsigned NXCD = 8;
    if (nblocks < NXCD * 2) return b;
    unsigned per = nblocks / NXCD, rem = nblocks % NXCD;
    unsigned xcd = b % NXCD, k = b / NXCD;
    // XCDs < rem own per+1 tiles
    unsigned start = xcd * per + (xcd < rem ? xcd : rem);
    unsigned cnt = per + (xcd < rem ? 1u : 0u);
    if (k < cnt) return start + k;
    return b;  // unreachable when nblocks is covered exactly
}

// Index sizes the kernels assume: voxel and bin indices are 32-bit, and the
// forward projector forms a voxel index as (z * ny + y) * nx + x and the
// backprojector a bin index as ib * na + ia with 24-bit multiply-adds
// (mad24), so each factor must stay below 2^24.
inline bool index_fits(long nx, long ny, long nz, long na, long nb)
{
    constexpr long L24 = 1L << 24;
    return nx * ny * nz < (1L << 32) && na * nb < (1L << 31) && nx < L24 && ny * nz < L24 &&
           na < L24 && nb < L24;
}
#define ASD_INDEX_LIMITS                                                                  \
    "volume or detector too large (32-bit voxel and bin indices; nx, na, nb and ny*nz " \
    "below 2^24)"

// a * b + c for a, b < 2^24 (exact while the result is below 2^32): a
// full-rate 24-bit multiply-add instead of a quarter-rate 32-bit multiply.
// The empty asm keeps the result opaque: when only its low 24 bits feed
// another mad24, the optimiser would otherwise rewrite this one as a 32-bit
// multiply (same low bits, four times the cost).
__host__ __device__ __forceinline__ unsigned mad24(unsigned a, unsigned b, unsigned c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    unsigned p = __umul24(a, b) + c;
    asm("" : "+v"(p));
    return p;
#else
    return a * b + c;
#endif
}

// Sum over each group of W consecutive lanes (W a power of two, at most 64),
// returned in every lane of the group.  Within a 16-lane row the partners
// are exchanged by DPP lane moves fused into the adds (quad swaps, then the
// half-row and row mirrors), which cost no more than the adds themselves;
// only the exchanges across rows go through ds_bpermute.
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}
template <int W, class R>
__device__ __forceinline__ R group_sum(R v)
{
    static_assert(W >= 1 && W <= 64 && (W & (W - 1)) == 0, "lane groups of 2^k <= 64");
    constexpr int QUAD_XOR1 = 0xB1, QUAD_XOR2 = 0x4E, ROW_HALF_MIRROR = 0x141, ROW_MIRROR = 0x140;
    if constexpr (W >= 2) v += dpp_move<QUAD_XOR1>(v);
    if constexpr (W >= 4) v += dpp_move<QUAD_XOR2>(v);
    if constexpr (W >= 8) v += dpp_move<ROW_HALF_MIRROR>(v);
    if constexpr (W >= 16) v += dpp_move<ROW_MIRROR>(v);
    if constexpr (W >= 32) v += __shfl_xor(v, 16, 64);
    if constexpr (W >= 64) v += __shfl_xor(v, 32, 64);
    return v;
}

// Kernel timing by the dispatch itself.  While a LaunchTiming is installed on
// the calling thread, klaunch hands its events to hipExtLaunchKernel, which
// stamps `start` with the first kernel's start and `stop` with the end of
// the last one: the kernels' own execution time, without the launch and
// queue gaps that events recorded around a launch also measure.
struct LaunchTiming {
    hipEvent_t start, stop;
    int launches;
};
inline thread_local LaunchTiming *t_launch_timing = nullptr;

template <typename... KArgs, typename... Args>
inline void klaunch(void (*kernel)(KArgs...), dim3 grid, dim3 block, unsigned shmem,
                    hipStream_t s, Args... args)
{
    if (LaunchTiming *t = t_launch_timing) {
        hipExtLaunchKernelGGL(kernel, grid, block, shmem, s, t->launches == 0 ? t->start : nullptr,
                              t->stop, 0u, KArgs(args)...);
        t->launches++;
    } else {
        kernel<<<grid, block, shmem, s>>>(KArgs(args)...);
    }
}

}  // namespace asd
