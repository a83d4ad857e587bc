// The TV-gradient stencils of the reference (imgTVGradMin.cpp:84-97,
// imgTVGradMin3D.cpp:114-140), shared by the single-plan descent (tv.hip) and
// the batched one (batch.hip) so both evaluate the same expressions.
#pragma once

#include "asd_internal.hpp"

namespace asd {

// Image values as the stencil sees them: zero outside the image, and with
// STEPPED the previous sub-iteration's step applied on the fly,
// u_i = T(u_{i-1} - s * v_{i-1}), s = alpha * d / ||v_{i-1}|| (the value the
// unfused step kernel would have stored).
template <class T, class A, bool STEPPED>
struct TvSrc {
    const T *u, *v;
    double s;
    int nx, ny, nz;
    __device__ __forceinline__ T stored(size_t i) const
    {
#pragma clang fp contract(off)
        if constexpr (STEPPED) return T((double)u[i] - s * (double)v[i]);
        else return u[i];
    }
    __device__ __forceinline__ A operator()(int x, int y, int z) const
    {
        if (x < 0 || x >= nx || y < 0 || y >= ny || z < 0 || z >= nz) return A(0);
        return A(stored(((size_t)z * ny + y) * nx + x));
    }
};

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
    A g1 = ((p - rm) + (p - cm)) / sqrt(eps + (p - rm) * (p - rm) + (p - cm) * (p - cm));
    A g2 = -(rp - p) / sqrt(eps + (rp - p) * (rp - p) + (rp - rp_cm) * (rp - rp_cm));
    A g3 = -(cp - p) / sqrt(eps + (cp - p) * (cp - p) + (cp - rm_cp) * (cp - rm_cp));
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
           sqrt(eps + (p - xm) * (p - xm) + (p - ym) * (p - ym) + (p - zm) * (p - zm));
    A g2 = (A(2) * p - A(2) * zp) /
           sqrt(eps + (zp - xm_zp) * (zp - xm_zp) + (zp - ym_zp) * (zp - ym_zp) + (zp - p) * (zp - p));
    A g3 = (A(2) * p - A(2) * yp) /
           sqrt(eps + (yp - xm_yp) * (yp - xm_yp) + (yp - p) * (yp - p) + (yp - yp_zm) * (yp - yp_zm));
    A g4 = (A(2) * p - A(2) * xp) /
           sqrt(eps + (xp - p) * (xp - p) + (xp - xp_ym) * (xp - xp_ym) + (xp - xp_zm) * (xp - xp_zm));
    return g1 + g2 + g3 + g4;
}

}  // namespace asd
