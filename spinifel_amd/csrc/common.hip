// Error state and device queries of libmtip.
#include "mtip_internal.h"

namespace mtip {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg)
{
    set_error(msg);
    return code;
}

}  // namespace mtip

extern "C" int mtip_abi_version(void) { return MTIP_ABI_VERSION; }

extern "C" const char *mtip_last_error(void) { return mtip::g_last_error.c_str(); }

extern "C" int mtip_mem_info(int device, size_t *free_bytes, size_t *total_bytes)
{
    MTIP_REQUIRE(free_bytes && total_bytes, MTIP_ERR_ARG, "mtip_mem_info: null output");
    int cur = 0;
    MTIP_HIP_CHECK(hipGetDevice(&cur));
    MTIP_HIP_CHECK(hipSetDevice(device));
    hipError_t e = hipMemGetInfo(free_bytes, total_bytes);
    (void)hipSetDevice(cur);
    if (e != hipSuccess) return mtip::fail(MTIP_ERR_HIP, std::string("hipMemGetInfo: ") + hipGetErrorString(e));
    return MTIP_OK;
}
