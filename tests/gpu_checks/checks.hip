// GPU checks of arithmetic identities the kernels rely on (test
// infrastructure, built by tests/gpu_checks/build.py).
//
// check_sqrt_forms: on every float whose bits lie in [lo, hi), compares the
// library's square roots (sqrt_finite_ge_tiny, sqrt_ge_tiny) and, as a
// control that the check sees a wrong root, the bare hardware sqrt with the
// compiler's correctly rounded sqrtf (this file is compiled with
// -fhip-fp32-correctly-rounded-divide-sqrt) and counts the inputs where the
// bits differ.  The forms are correct only as far as the hardware's rsq
// behaves as measured, so the check runs on the GPU the library runs on.
#include <hip/hip_runtime.h>
#include <cstdint>

#include "asd_internal.hpp"

namespace {

__global__ void sqrt_forms_kernel(uint32_t lo, uint32_t hi, unsigned long long *cnt,
                                  uint32_t *first)
{
    const uint32_t stride = gridDim.x * blockDim.x;
    unsigned long long n0 = 0, n1 = 0, n2 = 0;
    uint32_t f0 = ~0u, f1 = ~0u;
    // hi <= 2^32 - 2^25 (host check), so b + stride never wraps
    for (uint32_t b = lo + blockIdx.x * blockDim.x + threadIdx.x; b < hi; b += stride) {
        const float x = __uint_as_float(b);
        const uint32_t ref = __float_as_uint(sqrtf(x));
        if (__float_as_uint(asd::sqrt_finite_ge_tiny(x)) != ref) {
            n0++;
            f0 = b < f0 ? b : f0;
        }
        if (__float_as_uint(asd::sqrt_ge_tiny(x)) != ref) {
            n1++;
            f1 = b < f1 ? b : f1;
        }
        n2 += __float_as_uint(__builtin_amdgcn_sqrtf(x)) != ref;
    }
    atomicAdd(&cnt[0], n0);
    atomicAdd(&cnt[1], n1);
    atomicAdd(&cnt[2], n2);
    atomicMin(&first[0], f0);
    atomicMin(&first[1], f1);
}

}  // namespace

// Returns 0 and fills mismatches[3] (finite form, form with fixup, bare
// hardware sqrt) and first[2] (smallest mismatching bits, 0xffffffff if none); -1 on a bad
// range or a HIP error.
extern "C" int check_sqrt_forms(uint32_t lo, uint32_t hi, unsigned long long *mismatches,
                                uint32_t *first)
{
    if (lo > hi || hi > 0xfe000000u) return -1;
    unsigned long long *d_cnt = nullptr;
    uint32_t *d_first = nullptr;
    int rc = -1;
    if (hipMalloc(&d_cnt, 3 * sizeof(unsigned long long)) != hipSuccess) return -1;
    if (hipMalloc(&d_first, 2 * sizeof(uint32_t)) != hipSuccess) {
        (void)hipFree(d_cnt);
        return -1;
    }
    if (hipMemset(d_cnt, 0, 3 * sizeof(unsigned long long)) == hipSuccess &&
        hipMemset(d_first, 0xff, 2 * sizeof(uint32_t)) == hipSuccess) {
        hipLaunchKernelGGL(sqrt_forms_kernel, dim3(65536), dim3(256), 0, 0, lo, hi, d_cnt,
                           d_first);
        if (hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
            hipMemcpy(mismatches, d_cnt, 3 * sizeof(unsigned long long),
                      hipMemcpyDeviceToHost) == hipSuccess &&
            hipMemcpy(first, d_first, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost) == hipSuccess)
            rc = 0;
    }
    (void)hipFree(d_cnt);
    (void)hipFree(d_first);
    return rc;
}
