// Storage formats of the measured sinogram P inside a plan.
//
//   precision 0  FP32
//   precision 1  16-bit fixed point per view: P = lo + step * q, q in [0, 65535],
//                lo / step from the view's range (error <= step / 2)
//   precision 2  FP16 (11-bit significand; kept for comparison)
//   precision 3  as 1 (the backprojector also sets up its rays in FP64)
//   precision 4  FP64, as MATLAB holds P (the reference-arithmetic plan)
//   precision 5  FP16, with plain FP32 sums (half the sinogram traffic of 0)
//
// Each codec reads one value of one view as FP64, the type the SART
// correction (P - vxl*y) is formed in.
#pragma once

#include <hip/hip_fp16.h>

#include <cstdint>

namespace asd {

struct SinoF32 {
    const float *p;
    __device__ __forceinline__ double at(long i) const { return (double)p[i]; }
};

struct SinoU16 {
    const uint16_t *p;
    double lo, step;
    __device__ __forceinline__ double at(long i) const { return lo + step * (double)p[i]; }
};

struct SinoF16 {
    const __half *p;
    __device__ __forceinline__ double at(long i) const { return (double)__half2float(p[i]); }
};

struct SinoF64 {
    const double *p;
    __device__ __forceinline__ double at(long i) const { return p[i]; }
};

inline size_t sino_bytes(int precision) { return precision == 0 ? 4 : precision == 4 ? 8 : 2; }

}  // namespace asd
