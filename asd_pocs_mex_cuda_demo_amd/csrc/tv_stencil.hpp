// The TV-gradient stencils of the reference (imgTVGradMin.cpp:84-97,
// imgTVGradMin3D.cpp:114-140), shared by the single-plan descent (tv.hip) and
// the batched one (batch.hip) so both evaluate the same expressions.
#pragma once

#include "asd_internal.hpp"

namespace asd {

// One normalised TV step, u - alpha * d * v / ||v|| (imgTVGradMin.cpp:104-107,
// imgTVGradMin3D.cpp:157-160, where d = ||I - I0|| is fixed for the call).
// FP64 images are stepped in the reference's order, u - ((alpha * d) * v) /
// ||v||, so the drop-in TV and the reference-arithmetic plan round like the
// reference; FP32 images take the product with the step size
// s = (alpha * d) / ||v|| in FP64 and round once.
struct TvStepVal {
    double s;         // (alpha * d) / ||v||
    double ad, nrm;   // alpha * d and ||v||
    template <class T>
    __device__ __forceinline__ T apply(T u, T v) const
    {
#pragma clang fp contract(off)
        if constexpr (sizeof(T) == sizeof(double)) return u - (ad * v) / nrm;
        else return T((double)u - s * (double)v);
    }
};

__device__ __forceinline__ TvStepVal tv_step_val(double alpha_d, double nrm)
{
#pragma clang fp contract(off)
    return TvStepVal{alpha_d / nrm, alpha_d, nrm};
}

// Image values as the stencil sees them: zero outside the image, and with
// STEPPED the previous sub-iteration's step applied on the fly,
// u_i = step(u_{i-1}, v_{i-1}) (the value the unfused step kernel would have
// stored).
template <class T, class A, bool STEPPED>
struct TvSrc {
    const T *u, *v;
    TvStepVal st;
    int nx, ny, nz;
    __device__ __forceinline__ T stored(size_t i) const
    {
        if constexpr (STEPPED) return st.apply<T>(u[i], v[i]);
        else return u[i];
    }
    __device__ __forceinline__ A operator()(int x, int y, int z) const
    {
        if (x < 0 || x >= nx || y < 0 || y >= ny || z < 0 || z >= nz) return A(0);
        return A(stored(((size_t)z * ny + y) * nx + x));
    }
};

// The same values addressed relative to one voxel (x, y, z) at linear index
// i: the stencils ask for fixed offsets from it, so after inlining every
// neighbour is i plus a constant multiple of the row and plane strides, and
// no point forms its own index from coordinates.  A point outside the image
// reads the voxel itself and is replaced by 0 afterwards: every load is
// unconditional, so all of a stencil's loads are in flight together instead
// of one bounds-checked branch (and one memory round trip) after another.
template <class T, class A, bool STEPPED>
struct TvAt {
    const TvSrc<T, A, STEPPED> &S;
    size_t i;
    int x, y, z;
    __device__ __forceinline__ A operator()(int xx, int yy, int zz) const
    {
        const bool ok = (xx >= 0) & (xx < S.nx) & (yy >= 0) & (yy < S.ny) & (zz >= 0) &
                        (zz < S.nz);
        const long off = (long)(xx - x) + (long)(yy - y) * (long)S.nx +
                         (long)(zz - z) * ((long)S.nx * (long)S.ny);
        const A val = A(S.stored(i + (ok ? off : 0L)));
        return ok ? val : A(0);
    }
};

// The stencil's points, gathered ahead: TvGather loads the raw (u, v) of
// every point the stencil asks for, in the order it asks, into a TvRaw; a
// later evaluation through TvReplay applies the step to them in the same
// order.  The loads can then be issued before the step size is known (the
// fused kernel folds the previous norm meanwhile).  The point indices are
// compile-time constants after inlining, so the values stay in registers.
template <class T, int K>
struct TvRaw {
    T u[K], v[K];
    bool ok[K];
};

template <class T, class A, int K>
struct TvGather {
    const T *u, *v;
    size_t i;
    int x, y, z, nx, ny, nz;
    TvRaw<T, K> &raw;
    mutable int k;
    __device__ __forceinline__ A operator()(int xx, int yy, int zz) const
    {
        const bool ok = (xx >= 0) & (xx < nx) & (yy >= 0) & (yy < ny) & (zz >= 0) & (zz < nz);
        const long off = (long)(xx - x) + (long)(yy - y) * (long)nx +
                         (long)(zz - z) * ((long)nx * (long)ny);
        const size_t j = i + (ok ? off : 0L);
        raw.u[k] = u[j];
        raw.v[k] = v[j];
        raw.ok[k] = ok;
        k++;
        return A(0);
    }
};

template <class T, class A, int K>
struct TvReplay {
    const TvRaw<T, K> &raw;
    TvStepVal st;
    mutable int k;
    __device__ __forceinline__ A operator()(int, int, int) const
    {
        const int j = k++;
        // the value TvSrc<T, A, true>::stored gives for the same point
        const A val = A(st.apply<T>(raw.u[j], raw.v[j]));
        return raw.ok[j] ? val : A(0);
    }
};

// Coordinates of consecutive linear indices i, i + step, ... (x fastest),
// advanced without a division per voxel.
struct VoxelCursor {
    int x, y, z, nx, ny;
    __host__ __device__ __forceinline__ VoxelCursor(size_t i, int nx_, int ny_) : nx(nx_), ny(ny_)
    {
        if (i <= 0xffffffffu) {  // 32-bit division where the index allows it
            const unsigned u = (unsigned)i, r = u / (unsigned)nx;
            x = (int)(u - r * (unsigned)nx);
            y = (int)(r % (unsigned)ny);
            z = (int)(r / (unsigned)ny);
        } else {
            x = (int)(i % (size_t)nx);
            const size_t r = i / (size_t)nx;
            y = (int)(r % (size_t)ny);
            z = (int)(r / (size_t)ny);
        }
    }
    __host__ __device__ __forceinline__ void advance(int step)
    {
        x += step;
        while (x >= nx) {
            x -= nx;
            if (++y == ny) {
                y = 0;
                ++z;
            }
        }
    }
};

// Square root of a stencil denominator eps + (sum of squares) >= eps = 1e-8.
// In FP32 that argument is far above 2^-96, in the range where sqrt_ge_tiny
// gives the same bits as sqrtf; FP64 roots are taken as they are.
__device__ __forceinline__ float tv_sqrt(float x) { return sqrt_ge_tiny(x); }
__device__ __forceinline__ double tv_sqrt(double x) { return sqrt(x); }

// 2-D stencil at row r (fastest) and column c (imgTVGradMin.cpp:84-97).
template <class A, class Src>
__device__ __forceinline__ A tv2_grad(const Src &U, int r, int c)
{
#pragma clang fp contract(off)
    const A eps = A(1e-8);
    A p = U(r, c, 0);
    A rm = U(r - 1, c, 0), cm = U(r, c - 1, 0);
    A rp = U(r + 1, c, 0), rp_cm = U(r + 1, c - 1, 0);
    A cp = U(r, c + 1, 0), rm_cp = U(r - 1, c + 1, 0);
    A g1 = ((p - rm) + (p - cm)) / tv_sqrt(eps + (p - rm) * (p - rm) + (p - cm) * (p - cm));
    A g2 = -(rp - p) / tv_sqrt(eps + (rp - p) * (rp - p) + (rp - rp_cm) * (rp - rp_cm));
    A g3 = -(cp - p) / tv_sqrt(eps + (cp - p) * (cp - p) + (cp - rm_cp) * (cp - rm_cp));
    return g1 + g2 + g3;
}

// 3-D stencil at (x,y,z), x fastest (imgTVGradMin3D.cpp:114-140).
template <class A, class Src>
__device__ __forceinline__ A tv3_grad(const Src &U, int x, int y, int z)
{
#pragma clang fp contract(off)
    const A eps = A(1e-8);
    A p = U(x, y, z);
    A xm = U(x - 1, y, z), ym = U(x, y - 1, z), zm = U(x, y, z - 1);
    A zp = U(x, y, z + 1), xm_zp = U(x - 1, y, z + 1), ym_zp = U(x, y - 1, z + 1);
    A yp = U(x, y + 1, z), xm_yp = U(x - 1, y + 1, z), yp_zm = U(x, y + 1, z - 1);
    A xp = U(x + 1, y, z), xp_ym = U(x + 1, y - 1, z), xp_zm = U(x + 1, y, z - 1);
    A g1 = (A(6) * p - A(2) * (xm + ym + zm)) /
           tv_sqrt(eps + (p - xm) * (p - xm) + (p - ym) * (p - ym) + (p - zm) * (p - zm));
    A g2 = (A(2) * p - A(2) * zp) /
           tv_sqrt(eps + (zp - xm_zp) * (zp - xm_zp) + (zp - ym_zp) * (zp - ym_zp) + (zp - p) * (zp - p));
    A g3 = (A(2) * p - A(2) * yp) /
           tv_sqrt(eps + (yp - xm_yp) * (yp - xm_yp) + (yp - p) * (yp - p) + (yp - yp_zm) * (yp - yp_zm));
    A g4 = (A(2) * p - A(2) * xp) /
           tv_sqrt(eps + (xp - p) * (xp - p) + (xp - xp_ym) * (xp - xp_ym) + (xp - xp_zm) * (xp - xp_zm));
    return g1 + g2 + g3 + g4;
}

}  // namespace asd
