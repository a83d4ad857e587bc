// TV descent and FP64 reduction launchers shared by the TV entry points and
// the ASD-POCS driver.
#pragma once

#include "asd_internal.hpp"

namespace asd {

// Scalar slots of a TvScratch (FP64 results of grid reductions on the device).
enum : int {
    TV_SLOT_D2 = 0,     // ||I - I0||^2 of the TV call
    TV_SLOT_NRM2 = 1,   // ||v||^2 of the current sub-iteration
    TV_SLOT_DP2 = 2,    // driver: ||m - m_pre||^2
    TV_SLOT_DD2 = 3,    // driver: data error of the sweep
    TV_SLOT_RMSE2 = 4,  // driver: ||m - I0_truth||^2
    TV_SLOT_DG2 = 5,    // driver: ||m - f0||^2
    TV_NSLOTS = 8,
};

struct TvScratch {
    DeviceBuf buf;
    double *partials = nullptr;  // RED_MAX_SUMS * RED_MAX_BLOCKS block partials
    double *scalars = nullptr;   // TV_NSLOTS results
    int init();
};

template <class T>
int launch_sqdiff(const T *a, const T *b, size_t n, TvScratch &s, int slot, hipStream_t st);

// nTVIter steps of I -= alpha*d*v/||v|| (then positivity when pos == 1), in
// place on u.  d^2 = ||u - img_old||^2 is computed first unless have_d2, in
// which case scalars[TV_SLOT_D2] already holds it.  alpha_dev, when not
// NULL, overrides alpha from device memory; a non-zero *skip (device flag,
// may be NULL) makes every kernel return at entry.
// v, u2, v2: n-element work buffers (u2/v2 hold every other sub-iteration:
// each step is fused into the next gradient).
// serial_norms: d^2 and every ||v||^2 are summed in index order
// (serial_sumsq_kernel), the order of the reference's loops, instead of
// from block partials.
template <class T, class A>
int tv_descent(T *u, const T *img_old, T *v, T *u2, T *v2, int nx, int ny, int nz, bool three_d,
               double alpha, const double *alpha_dev, int nTVIter, int pos, TvScratch &s,
               hipStream_t st, bool have_d2, const int *skip = nullptr, bool serial_norms = false);

// *out = sum over i of (a[i] - b[i])^2 (b NULL: zeros) in index order (one
// wave; see serial_sumsq_kernel); skipped when *skip != 0.
template <class T>
int launch_serial_sumsq(const T *a, const T *b, size_t n, double *out, hipStream_t st,
                        const int *skip);

}  // namespace asd
