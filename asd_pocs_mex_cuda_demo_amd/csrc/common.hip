// Error state and device buffer helpers of libasdpocs.
#include <cstdlib>

#include "asd_internal.hpp"

namespace asd {

namespace {
thread_local int tl_status = ASD_OK;
thread_local std::string tl_msg;
}  // namespace

void set_error(int code, const std::string &msg)
{
    tl_status = code;
    tl_msg = msg;
}

void clear_error()
{
    tl_status = ASD_OK;
    tl_msg.clear();
}

int last_status() { return tl_status; }

int DeviceBuf::ensure(size_t n)
{
    if (n <= bytes && p) return ASD_OK;
    release();
    size_t want = n ? n : 1;
    // Allocation failures can be forced from the host for error-path tests.
    const char *lim = std::getenv("ASDPOCS_FAIL_ALLOC_ABOVE");
    if (lim && want > (size_t)std::strtoull(lim, nullptr, 10)) {
        set_error(ASD_ERR_ALLOC, "device allocation of " + std::to_string(want) +
                                     " bytes refused (ASDPOCS_FAIL_ALLOC_ABOVE)");
        return ASD_ERR_ALLOC;
    }
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        set_error(ASD_ERR_ALLOC, std::string("hipMalloc(") + std::to_string(want) +
                                     "): " + hipGetErrorString(e));
        return ASD_ERR_ALLOC;
    }
    bytes = want;
    return ASD_OK;
}

void DeviceBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

}  // namespace asd

extern "C" {

int asd_last_status(void) { return asd::last_status(); }

const char *asd_last_error(void) { return asd::tl_msg.c_str(); }

}  // extern "C"
