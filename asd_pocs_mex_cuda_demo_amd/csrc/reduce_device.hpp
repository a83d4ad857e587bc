// Deterministic FP64 reductions: wave butterflies and one partial per block;
// the partials are folded in a fixed order, by a one-block kernel
// (fold_partials) or inside every block of the consumer kernel (fold_rows).
// No atomics: a single counter that every block of a large grid increments
// serialises (~20 ns per block), which cost more than the data pass itself.
// Results do not depend on scheduling and are bit-identical run to run.
#pragma once

#include "asd_internal.hpp"

namespace asd {

constexpr int RED_THREADS = 256;
constexpr int RED_MAX_BLOCKS = 2048;  // grid cap of the grid-stride reductions
constexpr int RED_MAX_SUMS = 4;       // rows of partials a scratch buffer holds

__device__ __forceinline__ double wave_sum(double v) { return group_sum<64>(v); }

// Sum of `val` over the block (256 threads); valid in thread 0.
__device__ __forceinline__ double block_sum(double val)
{
    __shared__ double sh[RED_THREADS / 64];
    val = wave_sum(val);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) sh[w] = val;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < RED_THREADS / 64; i++) s += sh[i];
    }
    __syncthreads();
    return s;
}

inline unsigned red_blocks(size_t n)
{
    size_t b = (n + RED_THREADS - 1) / RED_THREADS;
    if (b > RED_MAX_BLOCKS) b = RED_MAX_BLOCKS;
    if (b == 0) b = 1;
    return (unsigned)b;
}

}  // namespace asd

namespace asd {

// Store the block's K sums at partials[k * gridDim.x + blockIdx.x].
template <int K>
__device__ __forceinline__ void store_partials(const double (&val)[K], double *partials)
{
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = block_sum(val[k]);
        if (threadIdx.x == 0) partials[(size_t)k * gridDim.x + blockIdx.x] = s;
    }
}

// This thread's share of row[0..n): row[t], row[t + 256], ... added in index
// order.  The first RED_MAX_BLOCKS / RED_THREADS of them are loaded together,
// so a row of up to RED_MAX_BLOCKS partials costs one memory round trip.
__device__ __forceinline__ double thread_row_sum(const double *row, int n)
{
    constexpr int PER = RED_MAX_BLOCKS / RED_THREADS;
    double val[PER];
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const int i = (int)threadIdx.x + j * RED_THREADS;
        val[j] = i < n ? row[i] : 0.0;
    }
    double t = 0.0;
#pragma unroll
    for (int j = 0; j < PER; j++)
        if ((int)threadIdx.x + j * RED_THREADS < n) t += val[j];
    for (int i = (int)threadIdx.x + PER * RED_THREADS; i < n; i += RED_THREADS) t += row[i];
    return t;
}

// Fold K rows of nblocks partials (one block of RED_THREADS threads); the
// totals are valid in every thread.
template <int K>
__device__ __forceinline__ void fold_rows(const double *partials, int nblocks, double (&tot)[K])
{
    __shared__ double res[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        double t = thread_row_sum(partials + (size_t)k * nblocks, nblocks);
        t = block_sum(t);
        if (threadIdx.x == 0) res[k] = t;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) tot[k] = res[k];
}

// sum over i of (a[i] - b[i])^2 (b NULL: zeros) in index order, rounded
// exactly as a serial loop forms it (the FP64 norms of the reference's TV
// loops, imgTVGradMin.cpp:58-59, :96, imgTVGradMin3D.cpp:150-153): one wave
// loads SER_U * 64 consecutive values at a time and adds their squares lane
// by lane from wave-uniform reads.  Lanes past n add +0, which leaves the
// non-negative sum unchanged.  Latency-bound (one dependent FP64 add per
// element); the reference-arithmetic plan uses it where results must round
// like the reference's.  Launch with one block of 64 threads.
constexpr int SER_U = 8;

__device__ __forceinline__ double readlane_f64(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}

template <class T>
__global__ __launch_bounds__(64) void serial_sumsq_kernel(const T *__restrict__ a,
                                                          const T *__restrict__ b, size_t n,
                                                          double *out, const int *skip)
{
#pragma clang fp contract(off)
    if (skip && *skip) return;
    const int lane = threadIdx.x;
    double acc = 0.0;
    for (size_t base = 0; base < n; base += 64 * SER_U) {
        double sq[SER_U];
#pragma unroll
        for (int u = 0; u < SER_U; u++) {
            const size_t i = base + (size_t)u * 64 + lane;
            double e = 0.0;
            if (i < n) e = (double)a[i] - (b ? (double)b[i] : 0.0);
            sq[u] = e * e;
        }
#pragma unroll
        for (int u = 0; u < SER_U; u++)
#pragma unroll 8
            for (int l = 0; l < 64; l++) acc += readlane_f64(sq[u], l);
    }
    if (lane == 0) *out = acc;
}

// out[k] = sum of row k (one block); skipped when *skip != 0.
template <int K>
__global__ __launch_bounds__(RED_THREADS) void fold_partials(const double *partials, int nblocks,
                                                            double *out, const int *skip)
{
    if (skip && *skip) return;
    double tot[K];
    fold_rows<K>(partials, nblocks, tot);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; k++) out[k] = tot[k];
}

}  // namespace asd
