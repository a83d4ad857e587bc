// Device-side geometry of the projectors.  Each function evaluates the
// reference's FP32 expressions in the reference's order with contraction
// off, so the GPU reproduces its rounding (the FP32 chord end points lose
// ~1e-4 voxel to cancellation at far sources; a different evaluation order
// would move every length by that much).
#pragma once

#include "asd_internal.hpp"

namespace asd {

__host__ __device__ __forceinline__ float absf(float a) { return a > 0 ? a : -a; }

// ---------------------------------------------------------------------------
// Forward projector: one unit slab of a ray (src/AxAtx/Ax_cone_gpu.cu:376-571).
// In the slab the sorted transverse extent [ta,tb] and the sorted axial extent
// [za,zb] (sorted independently, as the reference does) are split at the cell
// crossings rt = (ct2-ta)/(tb-ta) and rz = (cz2-za)/(zb-za) (1 without a
// crossing).  The pieces [0,min], [min,max], [max,1] lie in cells (ct1,cz1),
// (ct2,cz1) or (ct1,cz2), and (ct2,cz2); each in-range piece adds w*l*x.
// This one rule covers the reference's sixteen labelled cases 11..44.
// ---------------------------------------------------------------------------
struct AxRay {
    float u1, t1s, s1, s2, dz, q, tmp;
    // rtmp = 1 / tmp correctly rounded; mk: tmp and q lie where the division
    // by tmp may be finished from rtmp (div_by_tmp)
    float rtmp;
    int mk;
    // fast_t / fast_z: the slab's crossing fractions may use div_cell (below)
    int fast_t, fast_z;
    int nu2, nt2, nz2, nt, nz;
    int zlo, zhi;    // the slices summed: [zlo, zhi) (the view's z-window)
    int xdriven;
    // flat: a one-slice volume and a ray in its z = 0 face (s2 = +-0, dz = 1),
    // where every piece lies in slice 0 with rz = 1; dz1: dz == 1, where
    // x / dz == x exactly.  Both only skip work, never change a bit.
    int flat, dz1;
};

// The source lies in the plane z = 0 (cu:362-367), so the reference's axial
// coordinate z1 + s2 * (u - u1) adds z1 = 0.  Leaving that addition out
// changes only the sign of a zero product, and the slab's axial position
// p + nz/2 (or p / dz + nz/2) rounds -0 and +0 to the same value, so no
// result bit depends on it.

// x / tmp for one slab of the ray (the length factor of cu:402), finished
// from the ray's correctly rounded reciprocal by Markstein's correction:
// with y = RN(1/b) and q0 = RN(a*y) within one ulp of a/b, the residual
// r = a - b*q0 is exact in an FMA and RN(q0 + r*y) = RN(a/b) (P. Markstein,
// IBM J. Res. Dev. 34(1), 1990; Muller et al., Handbook of Floating-Point
// Arithmetic, the FMA-based division theorems).  The result is the IEEE
// quotient bit for bit whenever no step underflows or overflows, which
// mk guarantees (2^-60 <= tmp <= q <= 2^60 and 1 <= x / q <= ~2); other rays
// divide.
__host__ __device__ __forceinline__ float div_by_tmp(const AxRay &r, float x)
{
#pragma clang fp contract(off)
    if (!r.mk) return x / r.tmp;
    const float q0 = x * r.rtmp;
    const float res = fmaf(-r.tmp, q0, x);
    return fmaf(res, r.rtmp, q0);
}

// sqrt(x / tmp), the slab length of cu:402: the argument is >= 1, and
// finite where mk holds; other rays keep the fixup for +inf.
__host__ __device__ __forceinline__ float slab_root(const AxRay &r, float x)
{
    const float v = div_by_tmp(r, x);
    return r.mk ? sqrt_finite_ge_tiny(v) : sqrt_ge_tiny(v);
}

// (c - a) / (b - a) for a slab's cell crossing, bit for bit the IEEE
// quotient: the compiler's expansion of the FP32 division without its
// operand scaling and special-value fixup, which leave the result unchanged
// when no step under- or overflows.  That holds for these quotients when the
// cell grid has at least two cells along the axis (so the position of a
// slab end is zero or at least 2^-24 away from an integer, by Sterbenz's
// lemma on p + n/2 with n/2 >= 1) and the coordinates stay below 2^22
// (ax_ray_setup sets fast_t / fast_z only then): numerator and denominator
// lie in [2^-24, 2^23] and every residual is a normal number.
__host__ __device__ __forceinline__ float div_cell(float num, float den)
{
#if defined(__HIP_DEVICE_COMPILE__)
    float y = __builtin_amdgcn_rcpf(den);
    const float e = __builtin_fmaf(-den, y, 1.0f);
    y = __builtin_fmaf(e, y, y);
    float q = num * y;
    const float r0 = __builtin_fmaf(-den, q, num);
    q = __builtin_fmaf(r0, y, q);
    const float r1 = __builtin_fmaf(-den, q, num);
    return __builtin_fmaf(r1, y, q);
#else
    return num / den;
#endif
}

// The axial part of a ray's setup: everything that depends on the detector
// row's coordinate zdet.  du is the in-plane run along the driving axis
// (ax_ray_setup); rays of one detector column differ only in these fields.
__host__ __device__ __forceinline__ void ax_ray_axial(AxRay &r, float du, float zdet)
{
#pragma clang fp contract(off)
    const float z1 = 0.0f, z2 = zdet;
    r.q = r.tmp + (z1 - z2) * (z1 - z2);
    // a ray in the source's plane (every ray of a 2-D view) has s2 = +-0
    // without a division; the sign of that zero reaches no result (s2 >= 0
    // holds for both, and s2 * (u - u1) + nz/2 rounds both to the same value)
    r.s2 = z2 == z1 ? 0.0f : (z2 - z1) / du;
    r.mk = r.tmp >= 0x1p-60f && r.q <= 0x1p60f;
    r.flat = r.dz1 && r.nz == 1 && r.s2 == 0.0f;
}

// Setup of the ray from the source to detector point (ydet, zdet) (the
// reference's make_ray, cu:362-375); du receives the in-plane run that
// ax_ray_axial needs for further rows of the same column.
__host__ __device__ __forceinline__ AxRay ax_ray_setup(const ViewGeom &g, float ydet, float zdet,
                                                       float &du)
{
#pragma clang fp contract(off)
    AxRay r;
    float x1 = g.cos_phi * (-g.SO), y1 = g.sin_phi * (-g.SO);
    float x2 = g.cos_phi * g.OD - g.sin_phi * ydet;
    float y2 = g.sin_phi * g.OD + g.cos_phi * ydet;
    r.tmp = (x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2);
    r.xdriven = absf(x1 - x2) > absf(y1 - y2);
    r.u1 = r.xdriven ? x1 : y1;
    r.t1s = r.xdriven ? y1 : x1;
    du = r.xdriven ? (x2 - x1) : (y2 - y1);
    float dt = r.xdriven ? (y2 - y1) : (x2 - x1);
    r.s1 = dt / du;
    r.dz = g.dz;
    r.rtmp = 1.0f / r.tmp;
    r.nu2 = r.xdriven ? g.nx / 2 : g.ny / 2;
    r.nt2 = r.xdriven ? g.ny / 2 : g.nx / 2;
    r.nt = r.xdriven ? g.ny : g.nx;
    r.nz = g.nz;
    r.nz2 = g.nz / 2;
    r.zlo = g.z0;
    r.zhi = g.z0 + g.nzs;
    r.dz1 = g.dz == 1.0f;
    // coordinates below 2^22 and at least two cells along each axis (div_cell)
    const bool small = g.SO + g.OD + (float)g.nx + (float)g.ny + (float)g.nz < 0x1p21f;
    r.fast_t = small && g.nx >= 2 && g.ny >= 2;
    r.fast_z = small && g.nz >= 2;
    ax_ray_axial(r, du, zdet);
    return r;
}

__host__ __device__ __forceinline__ AxRay ax_ray_setup(const ViewGeom &g, float ydet, float zdet)
{
    float du;
    return ax_ray_setup(g, ydet, zdet, du);
}

// Detector coordinate of bin i of n at pitch det (voxels), as the MEX
// gateways tabulate it (Ax_main_test.cpp:156-160: FP64 arithmetic rounded
// to FP32); host and device give the same value.
__host__ __device__ __forceinline__ float det_coord(int i, int n, float det)
{
#pragma clang fp contract(off)
    return (float)(((double)i - (double)n / 2.0 + 0.5) * (double)det);
}

// Derived per-view constants (host and device give the same FP32 values).
// Footprint cutting stays off until view_tight has checked the view's rays.
__host__ __device__ __forceinline__ void geom_finish(ViewGeom &g)
{
#pragma clang fp contract(off)
    g.fp_d = sqrtf((1 + g.dz * g.dz) / 2);
    // corner offsets of a voxel in the view frame (x' = c x + s y, y' = -s x + c y)
    for (int k = 0; k < 4; k++) {
        const float sx = (k & 1) ? 0.5f : -0.5f, sy = (k & 2) ? 0.5f : -0.5f;
        g.fp_cx[k] = sx * g.cos_phi + sy * g.sin_phi;
        g.fp_cy[k] = sy * g.cos_phi - sx * g.sin_phi;
    }
    g.fp_rdy = 1.0f / g.dy_det;
    g.fp_rdz = 1.0f / g.dz_det;
    g.tight = 0;
    g.general = 0;
    g.div_nx = make_fastdiv((unsigned)g.nx);
    g.div_ny = make_fastdiv((unsigned)g.ny);
    g.div_na = make_fastdiv((unsigned)g.na);
    g.z0 = 0;
    g.nzs = g.nz;
}

// Whether the backprojector footprint may be cut to the voxel's shadow: no
// ray of the view takes find_l_3d's a == 0 / b == 0 fallbacks (cu:131-135),
// which ignore a coordinate and give non-zero lengths outside the shadow.
// a and b are formed exactly as make_ray_rec forms them.
inline int view_tight(const ViewGeom &g, const float *y_det)
{
#pragma clang fp contract(off)
    const float x1 = g.cos_phi * (-g.SO), y1 = g.sin_phi * (-g.SO);
    for (int ia = 0; ia < g.na; ia++) {
        const float x2 = g.cos_phi * g.OD - g.sin_phi * y_det[ia];
        const float y2 = g.sin_phi * g.OD + g.cos_phi * y_det[ia];
        if (x2 - x1 == 0.0f || y2 - y1 == 0.0f) return 0;
    }
    return 1;
}

// Whether make_ray_rec gives every ray of the view a general mode with exact
// face offsets (RAY_X_FIRST or RAY_Y_FIRST), so the backprojector may call
// chord3_sides without dispatching on the mode: the view is tight (no a == 0
// or b == 0), dz == 1, no ray is flat (c == 0 in a one-slice volume), and no
// slope is a nonzero value below 2^-125.  The last is shown from bounds: every
// slope is a quotient of two of the nonzero |a|, |b|, |c| (a zero c gives the
// slopes 0 and +-inf), so it is at least min / max of those, rounded
// monotonically.  A view the bounds cannot clear keeps the dispatching walk.
// a, b and c are formed exactly as make_ray_rec forms them.
inline int view_general(const ViewGeom &g, const float *y_det, const float *z_det)
{
#pragma clang fp contract(off)
    if (!g.tight || g.dz != 1.0f) return 0;
    const float x1 = g.cos_phi * (-g.SO), y1 = g.sin_phi * (-g.SO), z1 = 0.0f;
    float lo = 3.0e38f, hi = 0.0f;
    for (int ia = 0; ia < g.na; ia++) {
        const float x2 = g.cos_phi * g.OD - g.sin_phi * y_det[ia];
        const float y2 = g.sin_phi * g.OD + g.cos_phi * y_det[ia];
        const float a = absf(x2 - x1), b = absf(y2 - y1);
        lo = a < lo ? a : lo;
        lo = b < lo ? b : lo;
        hi = a > hi ? a : hi;
        hi = b > hi ? b : hi;
    }
    for (int ib = 0; ib < g.nb; ib++) {
        const float c = absf(z_det[ib] - z1);
        if (c == 0.0f) {
            if (g.nz == 1) return 0;
            continue;
        }
        if (!(c == c)) return 0;
        lo = c < lo ? c : lo;
        hi = c > hi ? c : hi;
    }
    return lo > 0.0f && (double)lo / (double)hi >= 0x1p-120;
}

// Reciprocal for conservative bounds only (one ulp on the device).
__host__ __device__ __forceinline__ float bound_rcp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// Running FP32 sum; with KAHAN the rounding error of every addition is
// carried in `c` (compensated summation; needs no reassociation, which the
// build guarantees with -ffp-contract=off and no fast-math).
template <bool KAHAN>
struct Acc {
    float s = 0.0f, c = 0.0f;
    __host__ __device__ __forceinline__ void add(float x)
    {
#pragma clang fp contract(off)
        if constexpr (KAHAN) {
            float y = x - c;
            float t = s + y;
            c = (t - s) - y;
            s = t;
        } else {
            s += x;
        }
    }
    // add x where ok holds, without a branch.  Uncompensated: +0 where ok
    // fails, which leaves the sum's bits unchanged (a sum that starts at +0
    // is never -0, and s + 0 == s for every other s).  Compensated: the sum
    // and carry are kept where ok fails (adding 0 would apply the carry).
    __host__ __device__ __forceinline__ void add_if(bool ok, float x)
    {
#pragma clang fp contract(off)
        if constexpr (KAHAN) {
            const float s0 = s, c0 = c;
            add(x);
            s = ok ? s : s0;
            c = ok ? c : c0;
        } else {
            s += ok ? x : 0.0f;
        }
    }
    __host__ __device__ __forceinline__ float value() const { return KAHAN ? s - c : s; }
};

// Volume reads of the forward projector.  A piece is addressed in units of
// VolRead::unit bytes: (cz * ny + y) * nx + x scaled by the unit, split as
// z * (ny * nx * unit) + t * st + iu * su (x or y is the slab index iu along
// the driving axis and the other the transverse cell t).  Pieces outside the
// volume (ok false) carry an arbitrary address:
//  - buffer form (volume below 2^30 voxels, nx below 2^22): byte offsets
//    through a buffer resource whose range is the volume; the hardware
//    returns 0 for an offset past the range and a value inside it is read
//    but not used, so no address is selected or cleared;
//  - wide form: element indices through a 64-bit pointer, an unused piece
//    reading element 0.
// On the host both read the element only for a piece in the volume.
template <bool WIDE>
struct VolRead;

template <>
struct VolRead<false> {
    static constexpr unsigned unit = 4;
    __amdgpu_buffer_rsrc_t rsrc;
    const float *p;
    __host__ __device__ VolRead(const float *x, size_t n) : p(x)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        // dword3 0x00020000: 32-bit data format, the form untyped loads use
        rsrc = __builtin_amdgcn_make_buffer_rsrc((void *)x, (short)0, (int)(unsigned)(n * 4),
                                                 0x00020000);
#endif
    }
    __host__ __device__ __forceinline__ float operator()(unsigned off, bool ok) const
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, off, 0, 0));
#else
        return ok ? p[off / 4] : 0.0f;
#endif
    }
};

template <>
struct VolRead<true> {
    static constexpr unsigned unit = 1;
    const float *p;
    __host__ __device__ VolRead(const float *x, size_t) : p(x) {}
    __host__ __device__ __forceinline__ float operator()(unsigned off, bool ok) const
    {
        return p[ok ? off : 0u];
    }
};

// An address past any volume's buffer range: the hardware returns 0 for it
// without reading memory.  Slabs wholly outside the volume (a large share of
// a fan's slabs) use it so their loads fetch nothing.
constexpr unsigned VOL_NONE = 0xffffffffu;

// Whether a volume of n voxels with rows of nx can use the buffer form.
__host__ __device__ __forceinline__ bool vol_buffer_ok(size_t n, int nx)
{
    return n < (size_t(1) << 30) && nx < (1 << 22);
}

// Address strides of a ray's pieces in units of unit bytes: st per
// transverse cell, su per slab, zs per slice (zs = nxu * ny with
// nxu = nx * unit, formed as two 24-bit products per piece).
struct SlabAddr {
    unsigned st, su, nxu;
};

__host__ __device__ __forceinline__ SlabAddr slab_addr(const AxRay &r, int nx, unsigned unit)
{
    const unsigned nxu = (unsigned)nx * unit;
    return SlabAddr{r.xdriven ? nxu : unit, r.xdriven ? unit : nxu, nxu};
}

// The (up to) three pieces of slab iu of ray r: address (see VolRead), weight
// w*l and whether the piece lies in the volume.  A piece outside the volume
// has an unspecified weight; its address is VOL_NONE when the whole slab lies
// outside, else arbitrary.
struct SlabW {
    unsigned off[3];
    float wl[3];
    bool ok[3];
};

// An unspecified value (no instruction): what an unused piece carries.
template <class T>
__host__ __device__ __forceinline__ T unspecified(T x)
{
    return __builtin_nondeterministic_value(x);
}

// The transverse half of slab iu of a ray: its sorted transverse extent
// [ta, tb], cells, crossing fraction and the transverse part of its pieces'
// addresses.  It depends only on the ray's in-plane line, which every ray of
// one detector column shares (rays that differ only in their detector row),
// so ax_rows_kernel computes it once for a bundle of rows.  FT: fast_t is
// known to hold (the caller has branched on it once for the whole march), so
// the crossing fraction is formed only by div_cell; otherwise both forms are
// evaluated and one is selected.
struct SlabT {
    float ua, ub;  // the slab's bounds along the driving axis
    float ta, tb;
    float d2p1;    // (tb - ta)^2 + 1, the in-plane factor of the slab length
    float rt;      // transverse crossing fraction (1 without a crossing)
    int ct1, ct2;
    bool in_t1, in_t2;
    unsigned tp1, tp2;  // ct * st + iu * su (meaningful for cells in the volume)
};

template <bool FT = false>
__host__ __device__ __forceinline__ SlabT ax_slab_t(const AxRay &r, const SlabAddr &A, int iu)
{
#pragma clang fp contract(off)
    SlabT t;
    t.ua = (float)(iu - r.nu2);
    t.ub = t.ua + 1;
    if (r.s1 >= 0) {
        t.ta = r.t1s + r.s1 * (t.ua - r.u1) + r.nt2;
        t.tb = r.t1s + r.s1 * (t.ub - r.u1) + r.nt2;
    } else {
        t.ta = r.t1s + r.s1 * (t.ub - r.u1) + r.nt2;
        t.tb = r.t1s + r.s1 * (t.ua - r.u1) + r.nt2;
    }
    t.ct1 = (int)floorf(t.ta);
    t.ct2 = (int)floorf(t.tb);
    t.in_t1 = t.ct1 >= 0 && t.ct1 < r.nt;
    t.in_t2 = t.ct2 >= 0 && t.ct2 < r.nt;
    const float d = t.tb - t.ta;
    t.d2p1 = d * d + 1;
    t.rt = 1.0f;
    if (t.ct1 != t.ct2 && (t.in_t1 || t.in_t2))
        t.rt = (FT || r.fast_t) ? div_cell((float)t.ct2 - t.ta, t.tb - t.ta)
                                : ((float)t.ct2 - t.ta) / (t.tb - t.ta);
    // 24-bit products (index_fits, vol_buffer_ok); exact for cells in the
    // volume, whose addresses lie below 2^32
    const unsigned ub = mad24((unsigned)iu, A.su, 0u);
    t.tp1 = mad24((unsigned)t.ct1, A.st, ub);
    t.tp2 = mad24((unsigned)t.ct2, A.st, ub);
    return t;
}

// What a march knows of its rays' flatness (AxRay::flat), fixed once per
// march by one branch: FLAT_ANY, tested per slab; FLAT_ALL, every ray is flat
// (a 2-D view); FLAT_NONE, no ray is (a volume of more than one slice, or
// dz != 1, where flat is false by its definition).
enum { FLAT_ANY = 0, FLAT_ALL = 1, FLAT_NONE = 2 };

// No ray of the view is flat: AxRay::flat needs dz == 1 and a one-slice volume.
__host__ __device__ __forceinline__ bool view_has_no_flat_ray(const ViewGeom &g)
{
    return g.nz != 1 || g.dz != 1.0f;
}

// The rest of slab iu of ray r given its transverse half: the axial extent
// and cells, the slab length and the three pieces.  FLAT (above): with
// FLAT_ALL the axial cells are the constants of a flat ray, with FLAT_NONE
// they are always computed.
template <int FLAT = FLAT_ANY>
__host__ __device__ __forceinline__ void ax_slab_z(const AxRay &r, const SlabT &t,
                                                   const SlabAddr &A, int ny, SlabW &s)
{
#pragma clang fp contract(off)
    int cz1 = 0, cz2 = 0;
    float za = 0.0f, zb = 0.0f;
    if (FLAT == FLAT_NONE || (FLAT == FLAT_ANY && !r.flat)) {
        // (z1 + s2 * (u - u1)) / dz + nz/2 with z1 = 0 (see above); x / 1 == x,
        // so dz == 1 skips the division
        float pa, pb;
        if (r.s2 >= 0) {
            pa = r.s2 * (t.ua - r.u1);
            pb = r.s2 * (t.ub - r.u1);
        } else {
            pa = r.s2 * (t.ub - r.u1);
            pb = r.s2 * (t.ua - r.u1);
        }
        if (!r.dz1) {
            pa = pa / r.dz;
            pb = pb / r.dz;
        }
        za = pa + r.nz2;
        zb = pb + r.nz2;
        cz1 = (int)floorf(za);
        cz2 = (int)floorf(zb);
    }
    // in the volume's z-window (the whole volume, or a slab's slices)
    const bool in_z1 = cz1 >= r.zlo && cz1 < r.zhi, in_z2 = cz2 >= r.zlo && cz2 < r.zhi;
    s.ok[0] = t.in_t1 && in_z1;
    s.ok[2] = t.in_t2 && in_z2;
    // a slab wholly outside the volume has no piece in it.  (Forming its
    // pieces anyway, with only the addresses replaced by VOL_NONE, was
    // measured: 256^3 views 97.5 -> 96.1 us, 512^3 views 750 -> 771 us.)
    if (!((t.in_t1 || t.in_t2) && (in_z1 || in_z2))) {
        s.ok[1] = false;
        for (int p = 0; p < 3; p++) {
            s.off[p] = VOL_NONE;
            s.wl[p] = unspecified(0.0f);
        }
        return;
    }
    float rz = 1.0f;
    if (cz1 != cz2) rz = r.fast_z ? div_cell((float)cz2 - za, zb - za) : ((float)cz2 - za) / (zb - za);
    const bool tfirst = !(t.rt > rz);  // the t crossing comes first (ties: t)
    s.ok[1] = (tfirst && t.in_t2 && in_z1) || (!tfirst && t.in_t1 && in_z2);
    // (d*d + 1) * q >= q >= tmp, so the root's argument is >= 1 (rounding is
    // monotone); where div_by_tmp uses the reciprocal (mk: q <= 2^60,
    // tmp >= 2^-60, (d*d + 1) <= ~2) it is also finite
    const float l = slab_root(r, t.d2p1 * r.q);
    const float lo = tfirst ? t.rt : rz, hi = tfirst ? rz : t.rt;
    // slice parts (cz - zlo) * ny * nx * unit, relative to the window's first
    // slice, in two 24-bit products (index_fits: cz * ny < 2^24;
    // vol_buffer_ok or index_fits: nx * unit < 2^24); a slice outside the
    // window gives an address of a piece that is not used
    const unsigned zp1 = mad24(mad24((unsigned)(cz1 - r.zlo), (unsigned)ny, 0u), A.nxu, 0u);
    const unsigned zp2 = mad24(mad24((unsigned)(cz2 - r.zlo), (unsigned)ny, 0u), A.nxu, 0u);
    s.off[0] = t.tp1 + zp1;
    s.off[1] = tfirst ? t.tp2 + zp1 : t.tp1 + zp2;
    s.off[2] = t.tp2 + zp2;
    s.wl[0] = lo * l;
    s.wl[1] = (hi - lo) * l;
    s.wl[2] = (1.0f - hi) * l;
}

template <bool FT = false, int FLAT = FLAT_ANY>
__host__ __device__ __forceinline__ void ax_slab_geom(const AxRay &r, const SlabAddr &A, int iu,
                                                      int ny, SlabW &s)
{
    ax_slab_z<FLAT>(r, ax_slab_t<FT>(r, A, iu), A, ny, s);
}

// The slabs [lo, hi] of a bundle of K rays that share their in-plane line
// (rays of one detector column, or K = 1) outside which no slab has a piece
// in the volume's window: a slab's pieces lie in cells floor(t) and
// floor(z) of its end points' transverse and axial coordinates
// (ax_slab_t / ax_slab_z), so a slab with a piece in the window has an end
// point with 0 <= t < nt and, for one of the rays, zlo <= z < zhi.  Those
// coordinates are linear in the slab position u; the u where they hold is
// solved in FP64 and widened by a margin far above the FP32 rounding of t and
// z (1e-2 cell plus 1e-5 of the coordinates' magnitudes), so no slab with a
// piece is left out.  Skipping the other slabs adds nothing to any sum (their
// pieces are never added), so the ray sums keep their bits.
// Reciprocal for the bound below (about 2^-26 relative on the device, where
// its error moves a bound by far less than the margin).
__host__ __device__ __forceinline__ double bound_rcp64(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcp(x);
#else
    return 1.0 / x;
#endif
}

__host__ __device__ __forceinline__ void ax_bundle_slabs(const AxRay *r, int K, int nu, int &lo,
                                                         int &hi)
{
    const double u1 = r[0].u1, nu2 = r[0].nu2;
    const double mag = fabs((double)r[0].u1) + (double)nu + 1.0;
    // u interval (in the reference's slab coordinate ua = iu - nu2) where the
    // transverse coordinate lies in [-dt, nt + dt)
    double uL = -1e30, uR = 1e30;
    {
        const double s1 = r[0].s1, t0 = (double)r[0].t1s + (double)r[0].nt2;
        const double dt = 1e-2 + 1e-5 * (fabs((double)r[0].t1s) + fabs(s1) * mag + r[0].nt);
        if (s1 == 0.0) {
            if (!(t0 >= -dt && t0 < r[0].nt + dt)) uL = 1e30, uR = -1e30;
        } else {
            const double rs = bound_rcp64(s1);
            const double a = u1 + (-dt - t0) * rs, b = u1 + (r[0].nt + dt - t0) * rs;
            uL = a < b ? a : b;
            uR = a < b ? b : a;
        }
    }
    // union over the rays of the u intervals where the axial coordinate lies
    // in [zlo - dz, zhi + dz)
    double zL = 1e30, zR = -1e30;
    for (int k = 0; k < K; k++) {
        const double s2 = r[k].s2, z0 = r[k].nz2;
        const double dz = 1e-2 + 1e-5 * (fabs(s2) * mag / r[k].dz + r[k].nz);
        double a, b;
        if (s2 == 0.0) {
            if (!(z0 >= r[k].zlo - dz && z0 < r[k].zhi + dz)) continue;
            a = -1e30;
            b = 1e30;
        } else {
            const double rs = r[k].dz * bound_rcp64(s2);
            a = u1 + (r[k].zlo - dz - z0) * rs;
            b = u1 + (r[k].zhi + dz - z0) * rs;
            if (a > b) {
                const double t = a;
                a = b;
                b = t;
            }
        }
        zL = a < zL ? a : zL;
        zR = b > zR ? b : zR;
    }
    uL = zL > uL ? zL : uL;
    uR = zR < uR ? zR : uR;
    // slab iu spans u in [iu - nu2, iu - nu2 + 1]; one more slab each side
    const double fl = floor(uL + nu2) - 2.0, fh = floor(uR + nu2) + 1.0;
    lo = fl < 0.0 ? 0 : fl > (double)nu ? nu : (int)fl;
    hi = fh < -1.0 ? -1 : fh > (double)(nu - 1) ? nu - 1 : (int)fh;
}

// U consecutive slabs of one lane (iu0, iu0 + stride, ...): all geometry
// first, then every load, then the sums in slab order.
template <int U, bool KAHAN, bool FT, class Vol, int FLAT = FLAT_ANY>
__host__ __device__ __forceinline__ void ax_slabs(const AxRay &r, const SlabAddr &A, const Vol &V,
                                                  int iu0, int stride, int nu, int ny,
                                                  Acc<KAHAN> &acc, Acc<KAHAN> &accl)
{
#pragma clang fp contract(off)
    SlabW s[U];
    float xv[U][3];
#pragma unroll
    for (int k = 0; k < U; k++) {
        const int iu = iu0 + k * stride;
        if (iu < nu) {
            ax_slab_geom<FT, FLAT>(r, A, iu, ny, s[k]);
        } else {
            for (int p = 0; p < 3; p++) {
                s[k].ok[p] = false;
                s[k].off[p] = VOL_NONE;
                s[k].wl[p] = unspecified(0.0f);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < U; k++)
#pragma unroll
        for (int p = 0; p < 3; p++) xv[k][p] = V(s[k].off[p], s[k].ok[p]);
#pragma unroll
    for (int k = 0; k < U; k++)
#pragma unroll
        for (int p = 0; p < 3; p++)
            if (s[k].ok[p]) {
                acc.add(s[k].wl[p] * xv[k][p]);
                accl.add(s[k].wl[p]);
            }
}

// ---------------------------------------------------------------------------
// Backprojector: line ∩ voxel length (find_l_3d cu:120-209, find_l cu:16-117)
// evaluated from the per-ray slope record.  (px,py,pz) = source - voxel centre.
// ---------------------------------------------------------------------------

// The chord between the first two candidate hits in test order (find_l_3d's
// rule, cu:144-206).  The first hit is a select chain over the candidates;
// the second is the chain over "hit with an earlier hit": the chain
// conditions are prefix masks (scalar work per wave), so a candidate costs
// one vector select per coordinate and point, and nothing is stored through
// an index (which would land in scratch memory).
template <int K>
struct Cands {
    bool f[K];
    float x[K], y[K], z[K];
};

template <int K>
__host__ __device__ __forceinline__ bool chord_points(const Cands<K> &c, float &ex, float &ey,
                                                      float &ez)
{
#pragma clang fp contract(off)
    bool g[K];  // candidate k is a hit and an earlier one is too
    bool any = c.f[0];
    g[0] = false;
#pragma unroll
    for (int k = 1; k < K; k++) {
        g[k] = c.f[k] && any;
        any = any || c.f[k];
    }
    float p0x = c.x[K - 1], p0y = c.y[K - 1], p0z = c.z[K - 1];
    float p1x = c.x[K - 1], p1y = c.y[K - 1], p1z = c.z[K - 1];
#pragma unroll
    for (int k = K - 2; k >= 0; k--) {
        p0x = c.f[k] ? c.x[k] : p0x;
        p0y = c.f[k] ? c.y[k] : p0y;
        p0z = c.f[k] ? c.z[k] : p0z;
    }
#pragma unroll
    for (int k = K - 2; k >= 1; k--) {
        p1x = g[k] ? c.x[k] : p1x;
        p1y = g[k] ? c.y[k] : p1y;
        p1z = g[k] ? c.z[k] : p1z;
    }
    bool two = false;
#pragma unroll
    for (int k = 1; k < K; k++) two = two || g[k];
    ex = p1x - p0x;
    ey = p1y - p0y;
    ez = p1z - p0z;
    return two;
}

// sqrtf of a squared chord: the same bits as sqrtf, by sqrt_ge_tiny where
// the argument is at least 2^-96; a shorter chord (two hits within 2^-48
// voxel) takes sqrtf.
__host__ __device__ __forceinline__ float chord_sqrt(float x)
{
    return x >= 0x1p-96f ? sqrt_ge_tiny(x) : sqrtf(x);
}

// General 2-D chord (find_l with a' != 0, b' != 0) in plane (u,w):
// pu = u1_0 - uc, pw = w1_0 - wc, s_wu = b'/a', s_uw = a'/b', and the face
// offsets ow = s_wu * hu, ou = s_uw * hw (per ray).
// Candidate list of a 2-D chord in test order (u faces first when U_FIRST).
template <bool U_FIRST>
__host__ __device__ __forceinline__ bool chord2_points(float cw, float ow, float cu, float ou,
                                                       float hu, float hw, float &ex, float &ey)
{
#pragma clang fp contract(off)
    const bool um = absf(cw - ow) <= hw, up = absf(cw + ow) <= hw;
    const bool wm = absf(cu - ou) <= hu, wp = absf(cu + ou) <= hu;
    constexpr int iu = U_FIRST ? 0 : 2, iw = U_FIRST ? 2 : 0;
    Cands<4> c;
    c.f[iu] = um; c.x[iu] = -hu; c.y[iu] = cw - ow; c.z[iu] = 0.0f;
    c.f[iu + 1] = up; c.x[iu + 1] = hu; c.y[iu + 1] = cw + ow; c.z[iu + 1] = 0.0f;
    c.f[iw] = wm; c.x[iw] = cu - ou; c.y[iw] = -hw; c.z[iw] = 0.0f;
    c.f[iw + 1] = wp; c.x[iw + 1] = cu + ou; c.y[iw + 1] = hw; c.z[iw + 1] = 0.0f;
    float ez;
    return chord_points<4>(c, ex, ey, ez);
}

__host__ __device__ __forceinline__ float chord2(float pu, float pw, float s_wu, float s_uw,
                                                 float ow, float ou, float hu, float hw,
                                                 bool u_first)
{
#pragma clang fp contract(off)
    const float cw = s_wu * (-pu) + pw;   // u faces: w coordinate
    const float cu = s_uw * (-pw) + pu;   // w faces: u coordinate
    float ex, ey;
    const bool two = u_first ? chord2_points<true>(cw, ow, cu, ou, hu, hw, ex, ey)
                             : chord2_points<false>(cw, ow, cu, ou, hu, hw, ex, ey);
    if (!two) return 0.0f;
    return chord_sqrt(ex * ex + ey * ey);
}

// The chord of find_l_3d from its six face candidates, given in face order
// (x-, x+, y-, y+, z-, z+): tested x faces then y faces (X_FIRST) or y faces
// then x faces, then the z faces.
template <bool X_FIRST>
__host__ __device__ __forceinline__ bool chord6(const bool (&f)[6], const float (&x)[6],
                                                const float (&y)[6], const float (&z)[6],
                                                float &ex, float &ey, float &ez)
{
    Cands<6> c;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int o = X_FIRST ? k : (k < 2 ? k + 2 : k < 4 ? k - 2 : k);
        c.f[k] = f[o];
        c.x[k] = x[o];
        c.y[k] = y[o];
        c.z[k] = z[o];
    }
    return chord_points<6>(c, ex, ey, ez);
}

// The six face candidates from the end coordinates c -+ o of each face pair
// (xy: the y coordinate on the x faces, ...), m = c - o and p = c + o.
struct FaceEnds {
    float xy_m, xy_p, xz_m, xz_p, yx_m, yx_p, yz_m, yz_p, zx_m, zx_p, zy_m, zy_p;
};

__host__ __device__ __forceinline__ bool chord3_points(bool x_first, const FaceEnds &E, float hx,
                                                       float hy, float hz, float &ex, float &ey,
                                                       float &ez)
{
#pragma clang fp contract(off)
    const bool f[6] = {absf(E.xy_m) <= hy && absf(E.xz_m) <= hz,
                       absf(E.xy_p) <= hy && absf(E.xz_p) <= hz,
                       absf(E.yx_m) <= hx && absf(E.yz_m) <= hz,
                       absf(E.yx_p) <= hx && absf(E.yz_p) <= hz,
                       absf(E.zx_m) <= hx && absf(E.zy_m) <= hy,
                       absf(E.zx_p) <= hx && absf(E.zy_p) <= hy};
    const float x[6] = {-hx, hx, E.yx_m, E.yx_p, E.zx_m, E.zx_p};
    const float y[6] = {E.xy_m, E.xy_p, -hy, hy, E.zy_m, E.zy_p};
    const float z[6] = {E.xz_m, E.xz_p, E.yz_m, E.yz_p, -hz, hz};
    return x_first ? chord6<true>(f, x, y, z, ex, ey, ez) : chord6<false>(f, x, y, z, ex, ey, ez);
}

// c -+ s * h as the reference forms it: the product o = s * h rounded, then
// c - o and c + o.
__host__ __device__ __forceinline__ void face_ends(float c, float s, float h, float &m, float &p)
{
#pragma clang fp contract(off)
    const float o = s * h;
    m = c - o;
    p = c + o;
}

// The same values when s * h is exact (h = 0.5 and RAY_HALF_EXACT): a fused
// multiply-add rounds c -+ s * h once, which is the rounding of c -+ o.
__host__ __device__ __forceinline__ void face_ends_fma(float c, float s, float h, float &m,
                                                       float &p)
{
    m = __builtin_fmaf(-h, s, c);
    p = __builtin_fmaf(h, s, c);
}

// find_l_3d's chord for a general ray (a, b != 0).  FMA: the face offsets
// s * h are exact (h = 0.5; RAY_X_FIRST / RAY_Y_FIRST), so each end point
// c -+ s*h is one fused multiply-add with the rounding of c -+ o.
template <bool FMA>
__host__ __device__ __forceinline__ float chord3_general(const RayRec &R, bool x_first, float px,
                                                         float py, float pz, float hz)
{
#pragma clang fp contract(off)
    const float hx = 0.5f, hy = 0.5f;
    const float xy_c = R.s_ba * (-px) + py;
    const float xz_c = R.s_ca * (-px) + pz;
    const float yx_c = R.s_ab * (-py) + px;
    const float yz_c = R.s_cb * (-py) + pz;
    const float zx_c = R.s_ac * (-pz) + px;
    const float zy_c = R.s_bc * (-pz) + py;
    FaceEnds E;
    auto ends = [](float c, float s, float h, float &m, float &p) {
        if constexpr (FMA)
            face_ends_fma(c, s, h, m, p);
        else
            face_ends(c, s, h, m, p);
    };
    ends(xy_c, R.s_ba, hx, E.xy_m, E.xy_p);
    ends(xz_c, R.s_ca, hx, E.xz_m, E.xz_p);
    ends(yx_c, R.s_ab, hy, E.yx_m, E.yx_p);
    ends(yz_c, R.s_cb, hy, E.yz_m, E.yz_p);
    ends(zx_c, R.s_ac, hz, E.zx_m, E.zx_p);
    ends(zy_c, R.s_bc, hz, E.zy_m, E.zy_p);
    float ex, ey, ez;
    const bool two = chord3_points(x_first, E, hx, hy, hz, ex, ey, ez);
    if (!two) return 0.0f;
    return chord_sqrt(ex * ex + ey * ey + ez * ez);
}

// The coordinate -sign(d_k) / 2 of the entry face of axis K (RayRec::side:
// the top byte of that FP32 value in byte 0, 1 or 3 for x, y or z, so each
// takes one instruction: a shift, a byte permute or a mask).
template <int K>
__host__ __device__ __forceinline__ float entry_half(unsigned side)
{
    static_assert(K >= 0 && K <= 2, "axis");
    if constexpr (K == 0) return __builtin_bit_cast(float, side << 24);
    if constexpr (K == 2) return __builtin_bit_cast(float, side & 0xff000000u);
#if defined(__HIP_DEVICE_COMPILE__)
    // result byte 3 = byte 1 of side, bytes 0..2 = 0 (selector 0x0c)
    return __builtin_bit_cast(float, __builtin_amdgcn_perm(0u, side, 0x010c0c0cu));
#else
    return __builtin_bit_cast(float, (side << 16) & 0xff000000u);
#endif
}

// find_l_3d's chord for a general ray with exact face offsets (RAY_X_FIRST,
// RAY_Y_FIRST), from the two sides of the voxel.  The six candidates of
// find_l_3d split into the three faces the ray enters through (E: the x, y
// or z face at -sign(d_k) / 2) and the three it leaves through (X).  The
// reference keeps the first two candidates that pass its tests, in its test
// order (cu:141-205), and returns their distance.  When exactly one E and
// one X candidate pass (a ray through the voxel away from its edges), those
// are the two, whatever the order, and the distance of two points does not
// depend on which comes first (p1 - p0 and p0 - p1 differ only in sign, so
// their squares are equal): the entry point is a select among three E
// candidates and the exit point among three X candidates.  Each candidate
// coordinate is the reference's c -+ s/2, formed as one fused multiply-add
// of the exact product s * (+-1/2), so every test and every coordinate has
// the reference's bits.  With fewer than two passes the length is 0, as in
// the reference.  The common cases run without a branch; every other
// pattern (several passes on one side, near an edge or a corner of the
// voxel) and chords shorter than 2^-48 (chord_sqrt) take the reference's
// ordered rule behind one branch.
__host__ __device__ __forceinline__ float chord3_sides(const RayRec &R, float px, float py,
                                                       float pz, float hz)
{
#pragma clang fp contract(off)
    const float hxe = entry_half<0>(R.side), hye = entry_half<1>(R.side),
                hze = entry_half<2>(R.side);
    const float xy_c = R.s_ba * (-px) + py;
    const float xz_c = R.s_ca * (-px) + pz;
    const float yx_c = R.s_ab * (-py) + px;
    const float yz_c = R.s_cb * (-py) + pz;
    const float zx_c = R.s_ac * (-pz) + px;
    const float zy_c = R.s_bc * (-pz) + py;
    // end points on the entry faces (c + s * hE) and the exit faces (c - s * hE)
    const float xy_e = __builtin_fmaf(hxe, R.s_ba, xy_c), xy_x = __builtin_fmaf(-hxe, R.s_ba, xy_c);
    const float xz_e = __builtin_fmaf(hxe, R.s_ca, xz_c), xz_x = __builtin_fmaf(-hxe, R.s_ca, xz_c);
    const float yx_e = __builtin_fmaf(hye, R.s_ab, yx_c), yx_x = __builtin_fmaf(-hye, R.s_ab, yx_c);
    const float yz_e = __builtin_fmaf(hye, R.s_cb, yz_c), yz_x = __builtin_fmaf(-hye, R.s_cb, yz_c);
    const float zx_e = __builtin_fmaf(hze, R.s_ac, zx_c), zx_x = __builtin_fmaf(-hze, R.s_ac, zx_c);
    const float zy_e = __builtin_fmaf(hze, R.s_bc, zy_c), zy_x = __builtin_fmaf(-hze, R.s_bc, zy_c);
    const float hx = 0.5f, hy = 0.5f;
    const bool fxe = absf(xy_e) <= hy && absf(xz_e) <= hz;
    const bool fye = absf(yx_e) <= hx && absf(yz_e) <= hz;
    const bool fze = absf(zx_e) <= hx && absf(zy_e) <= hy;
    const bool fxx = absf(xy_x) <= hy && absf(xz_x) <= hz;
    const bool fyx = absf(yx_x) <= hx && absf(yz_x) <= hz;
    const bool fzx = absf(zx_x) <= hx && absf(zy_x) <= hy;
    // at least two passes on a side (near an edge or a corner of the voxel),
    // at least one on each side, and exactly one on each side
    const bool many_e = (fxe && (fye || fze)) || (fye && fze);
    const bool many_x = (fxx && (fyx || fzx)) || (fyx && fzx);
    const bool two = (fxe || fye || fze) && (fxx || fyx || fzx) && !many_e && !many_x;
    const float ex = (fxx ? -hxe : fyx ? yx_x : zx_x) - (fxe ? hxe : fye ? yx_e : zx_e);
    const float ey = (fxx ? xy_x : fyx ? -hye : zy_x) - (fxe ? xy_e : fye ? hye : zy_e);
    const float ez = (fxx ? xz_x : fyx ? yz_x : -hze) - (fxe ? xz_e : fye ? yz_e : hze);
    const float q = ex * ex + ey * ey + ez * ez;
    // every lane takes the root; its value is used only for a chord, whose q
    // is finite (|e| <= 1 per axis) and, off the branch below, >= 2^-96 (where
    // there is no chord the argument may be anything: the operations on it
    // raise no exception and the result is discarded)
    const float l = sqrt_finite_ge_tiny(q);
    // several passes on a side, or a chord below 2^-48: the ordered rule.
    // Otherwise at most one pass in all, or one on each side only (two).
    if (many_e || many_x || (two && !(q >= 0x1p-96f)))
        return chord3_general<true>(R, R.mode == RAY_X_FIRST, px, py, pz, hz);
    return two ? l : 0.0f;
}

// FAN: for kernels that only see fan views, where every ray is flat or takes
// an a == 0 / b == 0 fallback: flat rays are tested first and general rays
// take the ordered rule (chord3_general), so the two-sided chord is not
// compiled in.  Same lengths either way.
template <bool FAN = false>
__host__ __device__ __forceinline__ float find_l_3d_rec(const RayRec &R, float px, float py, float pz,
                                               float hz, float dz)
{
#pragma clang fp contract(off)
    const float hx = 0.5f, hy = 0.5f;
    if constexpr (FAN) {
        if (R.mode >= RAY_FLAT_X)
            return chord2(px, py, R.s_ba, R.s_ab, R.s_ba * hx, R.s_ab * hy, hx, hy,
                          R.mode == RAY_FLAT_X);
        if (R.mode <= RAY_Y_FIRST)
            return chord3_general<true>(R, R.mode == RAY_X_FIRST, px, py, pz, hz);
    }
    // the general rays first, behind one test (the other modes are the
    // a == 0 / b == 0 fallbacks, which only views without a footprint cut
    // have, general rays whose face offsets are inexact products, and rays in
    // the face of a one-slice volume)
    if (R.mode <= RAY_Y_FIRST) return chord3_sides(R, px, py, pz, hz);
    // A flat ray (c == 0 in a one-slice volume, pz = -hz; every ray of a fan
    // view), next behind one test: s_ca = s_cb = +-0 put both hits' z at pz,
    // so every x / y face test's z condition holds (|pz| == hz) and ez = 0;
    // s_ac, s_bc = +-inf make the z-face tests NaN / inf comparisons that
    // fail.  What is left is the in-plane chord, evaluated by the same
    // expressions in the same order.
    if (R.mode >= RAY_FLAT_X)
        return chord2(px, py, R.s_ba, R.s_ab, R.s_ba * hx, R.s_ab * hy, hx, hy,
                      R.mode == RAY_FLAT_X);
    switch (R.mode) {
    case RAY_X_FIRST_P: return chord3_general<false>(R, true, px, py, pz, hz);
    case RAY_Y_FIRST_P: return chord3_general<false>(R, false, px, py, pz, hz);
    case RAY_A0_B0: return absf(py) <= hy ? dz : 0.0f;
    case RAY_A0_C0: return absf(pz) <= hz ? 1.0f : 0.0f;
    case RAY_A0_YZ_U: return chord2(py, pz, R.s_cb, R.s_bc, R.s_cb * hy, R.s_bc * hz, hy, hz, true);
    case RAY_A0_YZ_W:
        return chord2(py, pz, R.s_cb, R.s_bc, R.s_cb * hy, R.s_bc * hz, hy, hz, false);
    case RAY_B0_C0: return absf(pz) <= hz ? 1.0f : 0.0f;
    case RAY_B0_XZ_U: return chord2(px, pz, R.s_ca, R.s_ac, R.s_ca * hx, R.s_ac * hz, hx, hz, true);
    default: return chord2(px, pz, R.s_ca, R.s_ac, R.s_ca * hx, R.s_ac * hz, hx, hz, false);
    }
}

// Whether the in-plane line of a ray may meet the voxel's square: one of the
// four in-plane face conditions of find_l_3d holds within a margin of 1e-2
// voxel.  When none does, the line misses the square by more than every
// rounding difference between the face tests (~1e-4), so every face test of
// find_l_3d fails and the length is exactly 0: skipping such a (voxel, ray)
// pair leaves the sums bit-identical (adding 0 changes nothing).  The
// a == 0 / b == 0 fallbacks ignore a coordinate and are never skipped.
__host__ __device__ __forceinline__ bool inplane_may_hit(const RayRec &R, float px, float py)
{
#pragma clang fp contract(off)
    if (!ray_general(R.mode)) return true;
    const float lim = 0.5f + 1e-2f;
    const float xy_c = R.s_ba * (-px) + py, xy_o = R.s_ba * 0.5f;
    const float yx_c = R.s_ab * (-py) + px, yx_o = R.s_ab * 0.5f;
    return absf(xy_c - xy_o) <= lim || absf(xy_c + xy_o) <= lim || absf(yx_c - yx_o) <= lim ||
           absf(yx_c + yx_o) <= lim;
}

// Slope record of ray (ia, ib) of view g (the divisions of cu:142-196).
__host__ __device__ __forceinline__ RayRec make_ray_rec(const ViewGeom &g, float ydet, float zdet)
{
#pragma clang fp contract(off)
    float x1 = g.cos_phi * (-g.SO), y1 = g.sin_phi * (-g.SO), z1 = 0.0f;
    float x2 = g.cos_phi * g.OD - g.sin_phi * ydet;
    float y2 = g.sin_phi * g.OD + g.cos_phi * ydet;
    float z2 = zdet;
    float a = x2 - x1, b = y2 - y1, c = z2 - z1;
    RayRec R;
    R.s_ba = b / a; R.s_ab = a / b;
    R.s_ca = c / a; R.s_cb = c / b;
    R.s_ac = a / c; R.s_bc = b / c;
    R.side = 0;
    if (a == 0) {
        if (b == 0) R.mode = RAY_A0_B0;
        else if (c == 0) R.mode = RAY_A0_C0;
        else R.mode = absf(b) > absf(c) ? RAY_A0_YZ_U : RAY_A0_YZ_W;
    } else if (b == 0) {
        if (c == 0) R.mode = RAY_B0_C0;
        else R.mode = absf(a) > absf(c) ? RAY_B0_XZ_U : RAY_B0_XZ_W;
    } else {
        R.mode = absf(a) > absf(b) ? RAY_X_FIRST : RAY_Y_FIRST;
        // the face offsets s * 0.5 (and s * dz/2) are exact unless a product
        // is subnormal (|s| < 2^-125, s != 0) or dz != 1
        bool exact = g.dz == 1.0f;
        for (float s : {R.s_ba, R.s_ab, R.s_ca, R.s_cb, R.s_ac, R.s_bc})
            exact = exact && !(s != 0.0f && absf(s) < 0x1p-125f);
        if (c == 0 && g.nz == 1 && g.dz == 1.0f)
            R.mode += RAY_FLAT_X;
        else if (!exact)
            R.mode += RAY_X_FIRST_P;
        else  // entry faces at -sign(d)/2 (c == 0: no z face passes, either sign)
            R.side = (a > 0 ? 0xbfu : 0x3fu) | (b > 0 ? 0xbfu : 0x3fu) << 8 |
                     (c > 0 ? 0xbfu : 0x3fu) << 24;
    }
    return R;
}

// Detector footprint of voxel (ix,iy,iz) (cu:230-257): bins ia in [a_lo,a_hi],
// ib in [b_lo,b_hi] (clamped), plus the voxel centre.
struct Footprint {
    int a_lo, a_hi, b_lo, b_hi;
    float xc, yc, zc;
};

// Cuts the footprint to the bins whose ray passes within SHADOW_MARGIN
// voxel of the voxel: the shadow of its square through the source (the
// extreme projections of its four corners) and of its z extent over the
// square's depth, widened by the margin.  The reference's footprint is about
// twice as wide.  A ray outside the widened shadow misses the voxel by more
// than the margin.  find_l_3d's FP32 face tests credit a ray that misses by
// at most about 6e-5 voxel at a 1200-voxel source distance (rays aimed at
// voxel edges, measured against exact FP64 geometry), and the shadow itself
// is computed to about 1e-4, so with a margin of 1e-2 every test of such a
// ray fails and its length is exactly 0: the sums are unchanged
// (bit-identical when uncompensated).  Each 1e-2 of margin adds about 0.015
// bin per footprint side at 512^3 (0.05 kept 22% more pairs than the nonzero
// ones, 1e-2 keeps 8% more).
constexpr float SHADOW_MARGIN = 0.01f;

// The part of a voxel's footprint its whole column (ix, iy) shares: the
// in-plane bin range (reference bounds, clamped and cut to the square's
// shadow) and what the axial range needs from the column.
struct FootCol {
    float xc, yc;
    float tz;        // SD / ((x' + SO) * dz_det), the reference's axial scale (cu:247-253)
    int a_lo, a_hi;
    float m_near, m_far, eps;  // magnification range over the square, cut margin
    int cut;         // the axial cut applies (tight view, square in front of the source)
};

// Voxel centre coordinates in the reference's voxel frame (cu:243-245):
// i + 0.5 - n/2, which the reference forms in FP64 and rounds, is exact in
// FP32 for |i - n/2| < 2^22.
__host__ __device__ __forceinline__ float voxel_xc(const ViewGeom &g, int ix)
{
    return (float)(ix - g.nx / 2) + 0.5f;
}
__host__ __device__ __forceinline__ float voxel_yc(const ViewGeom &g, int iy)
{
    return (float)(iy - g.ny / 2) + 0.5f;
}
__host__ __device__ __forceinline__ float voxel_zc(const ViewGeom &g, int iz)
{
#pragma clang fp contract(off)
    return ((float)(iz - g.nz / 2) + 0.5f) * g.dz;
}

__host__ __device__ __forceinline__ FootCol footprint_col(const ViewGeom &g, int ix, int iy)
{
#pragma clang fp contract(off)
    FootCol c;
    const float SD = g.SO + g.OD;
    const int na2 = g.na / 2;
    c.yc = voxel_yc(g, iy);
    c.xc = voxel_xc(g, ix);
    // (xr, yr): the voxel centre in the view frame (x' towards the detector)
    const float xr = g.cos_phi * c.xc + g.sin_phi * c.yc;
    const float yr = -g.sin_phi * c.xc + g.cos_phi * c.yc;
    const float t = SD / ((xr + g.SO) * g.dy_det);
    const int a_hi = (int)floorf((yr + 1) * t + na2);
    const int a_lo = (int)floorf((yr - 1) * t + na2);
    c.tz = g.dz_det != g.dy_det ? SD / ((xr + g.SO) * g.dz_det) : t;
    c.a_lo = a_lo < 0 ? 0 : a_lo;
    c.a_hi = a_hi > g.na - 1 ? g.na - 1 : a_hi;
    c.cut = 0;
    c.m_near = c.m_far = c.eps = 0.0f;
    if (!g.tight) return c;
    // the detector coordinate y' * SD / (SO + x') of the square's four
    // corners (offsets fp_cx/fp_cy in the view frame); the corners also give
    // its nearest and farthest depth, hence the largest and smallest
    // magnification
    float umin = 3.0e38f, umax = -3.0e38f, m_near = 0.0f, m_far = 3.0e38f, dmin = 3.0e38f;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float depth = g.SO + (xr + g.fp_cx[k]);
        const float m = SD * bound_rcp(depth);
        const float u = (yr + g.fp_cy[k]) * m;
        umin = fminf(umin, u);
        umax = fmaxf(umax, u);
        m_near = fmaxf(m_near, m);
        m_far = fminf(m_far, m);
        dmin = fminf(dmin, depth);
    }
    if (dmin <= 0.0f) return c;  // the square reaches behind the source: keep all
    c.cut = 1;
    c.m_near = m_near;
    c.m_far = m_far;
    c.eps = SHADOW_MARGIN * m_near;
    // bins i with centre (i - n/2 + 1/2) * det inside [lo - eps, hi + eps]
    const float ia_off = 0.5f * (float)g.na - 0.5f;
    const int cut_lo = (int)ceilf((umin - c.eps) * g.fp_rdy + ia_off);
    const int cut_hi = (int)floorf((umax + c.eps) * g.fp_rdy + ia_off);
    c.a_lo = cut_lo > c.a_lo ? cut_lo : c.a_lo;
    c.a_hi = cut_hi < c.a_hi ? cut_hi : c.a_hi;
    return c;
}

// The footprint of voxel iz of a column: the axial bin range (reference
// bounds, clamped, and cut to the shadow of the voxel's z extent over the
// square's depth) with the column's in-plane range.
__host__ __device__ __forceinline__ Footprint footprint_z(const ViewGeom &g, const FootCol &c,
                                                          int iz)
{
#pragma clang fp contract(off)
    Footprint f;
    const int nb2 = g.nb / 2;
    const float d = g.fp_d;
    f.zc = voxel_zc(g, iz);
    f.yc = c.yc;
    f.xc = c.xc;
    const int b_hi = (int)floorf((f.zc + d) * c.tz + nb2);
    const int b_lo = (int)floorf((f.zc - d) * c.tz + nb2);
    f.a_lo = c.a_lo;
    f.a_hi = c.a_hi;
    f.b_lo = b_lo < 0 ? 0 : b_lo;
    f.b_hi = b_hi > g.nb - 1 ? g.nb - 1 : b_hi;
    if (c.cut) {
        const float ib_off = 0.5f * (float)g.nb - 0.5f, hz = 0.5f * g.dz;
        const float z0 = f.zc - hz, z1 = f.zc + hz;
        const float vmin = z0 * (z0 < 0.0f ? c.m_near : c.m_far);
        const float vmax = z1 * (z1 > 0.0f ? c.m_near : c.m_far);
        const int cut_lo = (int)ceilf((vmin - c.eps) * g.fp_rdz + ib_off);
        const int cut_hi = (int)floorf((vmax + c.eps) * g.fp_rdz + ib_off);
        f.b_lo = cut_lo > f.b_lo ? cut_lo : f.b_lo;
        f.b_hi = cut_hi < f.b_hi ? cut_hi : f.b_hi;
    }
    return f;
}

__host__ __device__ __forceinline__ Footprint footprint(const ViewGeom &g, int ix, int iy, int iz)
{
    return footprint_z(g, footprint_col(g, ix, iy), iz);
}

}  // namespace asd
