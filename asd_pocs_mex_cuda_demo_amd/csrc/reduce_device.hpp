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
