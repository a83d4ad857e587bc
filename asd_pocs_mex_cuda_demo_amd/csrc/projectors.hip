// Forward projector (ray-driven) and backprojector (voxel-driven) for gfx950,
// plus the drop-in host-pointer launchers that replace the reference's
// Ax_cone_mf_gpu_new / Atx_cone_mf_gpu_new (src/AxAtx/Ax_cone_gpu.cu:778, :299).
#include <mutex>

#include "projector_kernels.hpp"

namespace asd {

namespace {

struct EpiPlain {
    float *y, *sr;
    float scale;
    struct Pre {};
    __device__ __forceinline__ bool skip() const { return false; }
    __device__ __forceinline__ Pre load(long) const { return Pre{}; }
    __device__ __forceinline__ void operator()(long id, float acc, float accl, const Pre &) const
    {
#pragma clang fp contract(off)
        y[id] = acc * scale;
        sr[id] = accl * scale;
    }
};

struct EpiBackPlain {
    float *x, *sc;
    float scale;
    struct Pre {};
    __device__ __forceinline__ bool skip() const { return false; }
    __device__ __forceinline__ Pre load(long) const { return Pre{}; }
    __device__ __forceinline__ void operator()(long v, float acc, float accl, const Pre &) const
    {
#pragma clang fp contract(off)
        x[v] = acc * scale;
        sc[v] = accl * scale;
    }
};

__global__ __launch_bounds__(256) void ray_table_kernel(ViewGeom g,
                                                        const float *__restrict__ y_det,
                                                        const float *__restrict__ z_det,
                                                        RayRec *__restrict__ rt)
{
    long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long)g.na * g.nb) return;
    int ia = (int)(id % g.na), ib = (int)(id / g.na);
    rt[id] = make_ray_rec(g, y_det[ia], z_det[ib]);
}

}  // namespace

int ax_lanes_per_ray(int na, int nb, int nu)
{
    // At least 8 lanes per ray, and more until about 2^20 lanes work (4096
    // waves over 256 CUs), while keeping at least 2 slabs per lane; a 2-D
    // view (a few hundred rays) gets up to four waves per ray.  Eight lanes
    // on consecutive slabs of one ray read neighbouring voxels along the
    // driving axis and share the ray's setup over a quarter of its slabs;
    // one lane per ray, which a 1024 x 768 detector would otherwise get,
    // takes 1.9 times as long for a 512^3 view on MI355X.  (A 512-slab ray
    // at 128 lanes puts 1.5 blocks of a 768-ray view on each CU, which
    // leaves half the SIMDs with one wave more than the rest; at 256 lanes
    // every SIMD gets three.)
    long rays = (long)na * nb;
    int lpr = 1;
    while (lpr < 256 && (rays * lpr < (1L << 20) || lpr < 8) && nu / (2 * lpr) >= 2) lpr *= 2;
    return lpr;
}

AxShape ax_shape(int na, int nb, int nu)
{
    AxShape s;
    s.lpr = ax_lanes_per_ray(na, nb, nu);
    // Bundles of rows at the same 8 lanes share each slab's transverse
    // arithmetic among the bundle's rays (fewer waves, each longer); every
    // ray's sum is the same at 2 and 3 rows.  Measured per view on MI355X in
    // the plan: 256^3 with 384 x 384 bins 96.8 us with three rows against
    // 98.6 with two (two slabs per load batch); 512^3 with 1024 x 768 bins
    // 742 against 728 us, where two rows leave the chip four times as many
    // waves to schedule.  So views of at least 2^19 rays take two rows, the
    // others three.  Views with fewer rays (more than 8 lanes per ray) keep
    // single rays, and so do detectors of fewer than 16 rows, where the last
    // bundle's repeated rows would cost more.
    if (s.lpr == 8 && nb >= 16) s.rows = (long)na * nb >= (1L << 19) ? 2 : 3;
    return s;
}

int launch_ray_table(const ViewGeom &g, const float *y_det, const float *z_det, RayRec *rt,
                     hipStream_t s)
{
    long n = (long)g.na * g.nb;
    unsigned blocks = (unsigned)((n + 255) / 256);
    ray_table_kernel<<<blocks, 256, 0, s>>>(g, y_det, z_det, rt);
    ASD_LAUNCH_CHECK("ray_table_kernel");
    return ASD_OK;
}

int launch_ax(const ViewGeom &g, const float *x, float *y, float *sr, const float *y_det,
              const float *z_det, hipStream_t s)
{
    const int nu_max = g.nx > g.ny ? g.nx : g.ny;
    return dispatch_ax(ax_shape(g.na, g.nb, nu_max), g, x, y_det, z_det,
                       EpiPlain{y, sr, g.scale}, s);
}

int launch_atx(const ViewGeom &g, float *x, const float *y, float *sc, const float *y_det,
               const float *z_det, const RayRec *rt, hipStream_t s)
{
    return dispatch_atx_variant<false>(ATX_DIRECT, g, y, rt, EpiBackPlain{x, sc, g.scale}, s);
}

// ---------------------------------------------------------------------------
// Process-wide context of the drop-in launchers: one stream and grow-only
// device buffers reused across calls (the reference allocates and frees five
// buffers per call, cu:308-330 and :787-804).
// ---------------------------------------------------------------------------
namespace {

struct DropIn {
    std::mutex mu;
    hipStream_t stream = nullptr;
    DeviceBuf vol, vol2, proj, proj2, ydet, zdet, rt;
    int init()
    {
        if (!stream) ASD_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        return ASD_OK;
    }
};

DropIn &dropin()
{
    static DropIn *d = new DropIn();  // lives until process exit
    return *d;
}

int check_geom(const ViewGeom &g)
{
    ASD_ARG(g.nx > 0 && g.ny > 0 && g.nz > 0 && g.na > 0 && g.nb > 0,
            "sizes nx, ny, nz, na, nb must be positive");
    ASD_ARG(index_fits(g.nx, g.ny, g.nz, g.na, g.nb), ASD_INDEX_LIMITS);
    // distances and pitches in voxel units (the gateways' SO / vxl, det / vxl
    // and dz = 1); NaN fails every comparison
    ASD_ARG(g.SO > 0 && g.OD > 0 && g.SO < 3.0e38f && g.OD < 3.0e38f,
            "SO and OD must be positive and finite");
    ASD_ARG(g.dz > 0 && g.dy_det > 0 && g.dz_det > 0 && g.dz < 3.0e38f && g.dy_det < 3.0e38f &&
                g.dz_det < 3.0e38f,
            "dz and the detector pitches must be positive and finite");
    return ASD_OK;
}

ViewGeom make_geom(float cos_phi, float sin_phi, float SO, float OD, float scale,
                   float dy_det, float dz_det, float dz, int nx, int ny, int nz, int na, int nb)
{
    ViewGeom g;
    g.cos_phi = cos_phi; g.sin_phi = sin_phi; g.SO = SO; g.OD = OD; g.scale = scale;
    g.dz = dz; g.dy_det = dy_det; g.dz_det = dz_det;
    g.nx = nx; g.ny = ny; g.nz = nz; g.na = na; g.nb = nb;
    geom_finish(g);
    return g;
}

int ax_host(float *X, float *y, float *sr, const ViewGeom &g, const float *y_det,
            const float *z_det)
{
    int rc = check_geom(g);
    if (rc) return rc;
    ASD_ARG(X && y && sr && y_det && z_det, "NULL pointer argument");
    DropIn &d = dropin();
    std::lock_guard<std::mutex> lk(d.mu);
    if ((rc = d.init())) return rc;
    size_t N = (size_t)g.nx * g.ny * g.nz, M = (size_t)g.na * g.nb;
    if ((rc = d.vol.ensure(N * 4)) || (rc = d.proj.ensure(M * 4)) ||
        (rc = d.proj2.ensure(M * 4)) || (rc = d.ydet.ensure(g.na * 4)) ||
        (rc = d.zdet.ensure(g.nb * 4)))
        return rc;
    ASD_HIP(hipMemcpyAsync(d.vol.p, X, N * 4, hipMemcpyHostToDevice, d.stream));
    ASD_HIP(hipMemcpyAsync(d.ydet.p, y_det, g.na * 4, hipMemcpyHostToDevice, d.stream));
    ASD_HIP(hipMemcpyAsync(d.zdet.p, z_det, g.nb * 4, hipMemcpyHostToDevice, d.stream));
    if ((rc = launch_ax(g, d.vol.as<float>(), d.proj.as<float>(), d.proj2.as<float>(),
                        d.ydet.as<float>(), d.zdet.as<float>(), d.stream)))
        return rc;
    ASD_HIP(hipMemcpyAsync(y, d.proj.p, M * 4, hipMemcpyDeviceToHost, d.stream));
    ASD_HIP(hipMemcpyAsync(sr, d.proj2.p, M * 4, hipMemcpyDeviceToHost, d.stream));
    ASD_HIP(hipStreamSynchronize(d.stream));
    return ASD_OK;
}

int atx_host(float *X, const float *y, float *sc, const ViewGeom &g, const float *y_det,
             const float *z_det)
{
    int rc = check_geom(g);
    if (rc) return rc;
    ASD_ARG(X && y && sc && y_det && z_det, "NULL pointer argument");
    DropIn &d = dropin();
    std::lock_guard<std::mutex> lk(d.mu);
    if ((rc = d.init())) return rc;
    size_t N = (size_t)g.nx * g.ny * g.nz, M = (size_t)g.na * g.nb;
    if ((rc = d.vol.ensure(N * 4)) || (rc = d.vol2.ensure(N * 4)) ||
        (rc = d.proj.ensure(M * 4)) || (rc = d.ydet.ensure(g.na * 4)) ||
        (rc = d.zdet.ensure(g.nb * 4)) || (rc = d.rt.ensure(M * sizeof(RayRec))))
        return rc;
    ASD_HIP(hipMemcpyAsync(d.proj.p, y, M * 4, hipMemcpyHostToDevice, d.stream));
    ASD_HIP(hipMemcpyAsync(d.ydet.p, y_det, g.na * 4, hipMemcpyHostToDevice, d.stream));
    ASD_HIP(hipMemcpyAsync(d.zdet.p, z_det, g.nb * 4, hipMemcpyHostToDevice, d.stream));
    if ((rc = launch_ray_table(g, d.ydet.as<float>(), d.zdet.as<float>(), d.rt.as<RayRec>(),
                               d.stream)))
        return rc;
    ViewGeom gt = g;
    gt.tight = view_tight(g, y_det);  // host table: the rays' a and b can be checked here
    gt.general = view_general(gt, y_det, z_det);
    if ((rc = launch_atx(gt, d.vol.as<float>(), d.proj.as<float>(), d.vol2.as<float>(),
                         d.ydet.as<float>(), d.zdet.as<float>(), d.rt.as<RayRec>(), d.stream)))
        return rc;
    ASD_HIP(hipMemcpyAsync(X, d.vol.p, N * 4, hipMemcpyDeviceToHost, d.stream));
    ASD_HIP(hipMemcpyAsync(sc, d.vol2.p, N * 4, hipMemcpyDeviceToHost, d.stream));
    ASD_HIP(hipStreamSynchronize(d.stream));
    return ASD_OK;
}

// Device-pointer forms keep one ray table per thread of the caller.
thread_local DeviceBuf tl_rt;

// ---------------------------------------------------------------------------
// The backprojector's ray table for asd_atx_device_noalloc: a fixed device
// array instead of a library allocation, so that entry point allocates
// nothing.  Each device holds its own copy (32 MiB) until the process ends;
// every call on a device shares it and overwrites the first na * nb records
// (nothing in it needs to start at zero), so calls that overlap on one
// device (different streams) race.  In host code `scratch` names no device
// address: its address on the current device comes from
// hipGetSymbolAddress, once per device (scratch_table).
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_MAX_BINS = size_t(1) << 20;  // e.g. a 1024 x 1024 detector
static __device__ __align__(256) RayRec scratch[NOALLOC_MAX_BINS];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex scratch_mu;
RayRec *scratch_dev[NOALLOC_MAX_DEVICES] = {};

// The address of `scratch` on the current device (resolved on the first call
// for that device, with it current).
int scratch_table(RayRec **out)
{
    int dev = 0;
    ASD_HIP(hipGetDevice(&dev));
    ASD_ARG(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_dev[dev]) {
        void *ptr = nullptr;
        ASD_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        ASD_ARG(ptr, "hipGetSymbolAddress returned no address for the scratch table");
        scratch_dev[dev] = static_cast<RayRec *>(ptr);
    }
    *out = scratch_dev[dev];
    return ASD_OK;
}

}  // namespace
}  // namespace asd

using namespace asd;

extern "C" {

void Ax_cone_mf_gpu_new(float *X, float *y, float *sr, float cos_phi, float sin_phi,
                        float *y_det, float *z_det, float SO, float OD, float scale, float dz,
                        int nx, int ny, int nz, int na, int nb)
{
    clear_error();
    ViewGeom g = make_geom(cos_phi, sin_phi, SO, OD, scale, 1.0f, 1.0f, dz, nx, ny, nz, na, nb);
    ax_host(X, y, sr, g, y_det, z_det);
}

void Atx_cone_mf_gpu_new(float *X, float *y, float *sc, float cos_phi, float sin_phi,
                         float *y_det, float *z_det, float SO, float OD, float scale,
                         float dy_det, float dz_det, float dz, int nx, int ny, int nz, int na,
                         int nb)
{
    clear_error();
    ViewGeom g = make_geom(cos_phi, sin_phi, SO, OD, scale, dy_det, dz_det, dz, nx, ny, nz, na,
                           nb);
    atx_host(X, y, sc, g, y_det, z_det);
}

int asd_ax_device(const float *X, float *y, float *sr, float cos_phi, float sin_phi,
                  const float *y_det, const float *z_det, float SO, float OD, float scale,
                  float dz, int nx, int ny, int nz, int na, int nb, void *stream)
{
    clear_error();
    ViewGeom g = make_geom(cos_phi, sin_phi, SO, OD, scale, 1.0f, 1.0f, dz, nx, ny, nz, na, nb);
    int rc = check_geom(g);
    if (rc) return rc;
    ASD_ARG(X && y && sr && y_det && z_det, "NULL pointer argument");
    return launch_ax(g, X, y, sr, y_det, z_det, (hipStream_t)stream);
}

int asd_atx_device(float *X, const float *y, float *sc, float cos_phi, float sin_phi,
                   const float *y_det, const float *z_det, float SO, float OD, float scale,
                   float dy_det, float dz_det, float dz, int nx, int ny, int nz, int na, int nb,
                   void *stream)
{
    clear_error();
    ViewGeom g = make_geom(cos_phi, sin_phi, SO, OD, scale, dy_det, dz_det, dz, nx, ny, nz, na,
                           nb);
    int rc = check_geom(g);
    if (rc) return rc;
    ASD_ARG(X && y && sc && y_det && z_det, "NULL pointer argument");
    if ((rc = tl_rt.ensure((size_t)na * nb * sizeof(RayRec)))) return rc;
    hipStream_t s = (hipStream_t)stream;
    if ((rc = launch_ray_table(g, y_det, z_det, tl_rt.as<RayRec>(), s))) return rc;
    return launch_atx(g, X, y, sc, y_det, z_det, tl_rt.as<RayRec>(), s);
}

int asd_ax_slab_device(const float *X, float *y, float *sr, float cos_phi, float sin_phi,
                       const float *y_det, const float *z_det, float SO, float OD, float scale,
                       float dz, int nx, int ny, int nz, int na, int nb, int z0, int nzs,
                       void *stream)
{
    clear_error();
    ViewGeom g = make_geom(cos_phi, sin_phi, SO, OD, scale, 1.0f, 1.0f, dz, nx, ny, nz, na, nb);
    int rc = check_geom(g);
    if (rc) return rc;
    ASD_ARG(X && y && sr && y_det && z_det, "NULL pointer argument");
    ASD_ARG(z0 >= 0 && nzs > 0 && z0 <= nz - nzs, "the slab [z0, z0 + nzs) must lie in [0, nz)");
    g.z0 = z0;
    g.nzs = nzs;
    return launch_ax(g, X, y, sr, y_det, z_det, (hipStream_t)stream);
}

int asd_atx_slab_device(float *X, const float *y, float *sc, float cos_phi, float sin_phi,
                        const float *y_det, const float *z_det, float SO, float OD, float scale,
                        float dy_det, float dz_det, float dz, int nx, int ny, int nz, int na,
                        int nb, int z0, int nzs, void *stream)
{
    clear_error();
    ViewGeom g = make_geom(cos_phi, sin_phi, SO, OD, scale, dy_det, dz_det, dz, nx, ny, nz, na,
                           nb);
    int rc = check_geom(g);
    if (rc) return rc;
    ASD_ARG(X && y && sc && y_det && z_det, "NULL pointer argument");
    ASD_ARG(z0 >= 0 && nzs > 0 && z0 <= nz - nzs, "the slab [z0, z0 + nzs) must lie in [0, nz)");
    g.z0 = z0;
    g.nzs = nzs;
    if ((rc = tl_rt.ensure((size_t)na * nb * sizeof(RayRec)))) return rc;
    hipStream_t s = (hipStream_t)stream;
    if ((rc = launch_ray_table(g, y_det, z_det, tl_rt.as<RayRec>(), s))) return rc;
    return launch_atx(g, X, y, sc, y_det, z_det, tl_rt.as<RayRec>(), s);
}

int asd_atx_device_noalloc(float *X, const float *y, float *sc, float cos_phi, float sin_phi,
                           const float *y_det, const float *z_det, float SO, float OD,
                           float scale, float dy_det, float dz_det, float dz, int nx, int ny,
                           int nz, int na, int nb, void *stream)
{
    clear_error();
    ViewGeom g = make_geom(cos_phi, sin_phi, SO, OD, scale, dy_det, dz_det, dz, nx, ny, nz, na,
                           nb);
    int rc = check_geom(g);
    if (rc) return rc;
    ASD_ARG(X && y && sc && y_det && z_det, "NULL pointer argument");
    ASD_ARG((size_t)na * (size_t)nb <= NOALLOC_MAX_BINS,
            "asd_atx_device_noalloc: na * nb exceeds its fixed ray table of 1048576 bins "
            "(use asd_atx_device)");
    RayRec *rt = nullptr;
    if ((rc = scratch_table(&rt))) return rc;
    hipStream_t s = (hipStream_t)stream;
    if ((rc = launch_ray_table(g, y_det, z_det, rt, s))) return rc;
    return launch_atx(g, X, y, sc, y_det, z_det, rt, s);
}

}  // extern "C"
