// Allocation-free top-K matching: mtip_match_topk with its partial-key
// workspace in a device array inside the library instead of a caller buffer.
//
// `scratch` is one file-scope device array, one copy per GPU the library
// runs on, resident from the code object's load on that GPU until the
// process ends.  In host code its name is a host-side stand-in, not a device
// address: host code passes kernels only the address hipGetSymbolAddress
// returned for the current device, looked up once per device.  The GEMM
// kernels write every key slot the merge later reads, so nothing needs
// clearing between calls; calls that overlap on one GPU share the array and
// race.
#include <mutex>

#include "mtip_internal.h"

static __device__ __align__(256) uint64_t scratch[MTIP_STATIC_SCRATCH_BYTES / sizeof(uint64_t)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint64_t *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first
// call made with that device current).
int scratch_on_current_device(uint64_t **out)
{
    int dev = -1;
    MTIP_HIP_CHECK(hipGetDevice(&dev));
    MTIP_REQUIRE(dev >= 0 && dev < kMaxDevices, MTIP_ERR_ARG,
                 "mtip_match_topk_static: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        uint64_t *ptr = nullptr;
        MTIP_HIP_CHECK(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
        MTIP_REQUIRE(ptr != nullptr, MTIP_ERR_HIP, "mtip_match_topk_static: no scratch address");
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return MTIP_OK;
}

}  // namespace

extern "C" int mtip_match_topk_static(int precision, const void *slices, int64_t ld_s,
                                      const float *slice_scale, const double *slice_sumsq,
                                      int64_t N_o, const void *images, int64_t ld_d,
                                      const float *image_scale, int64_t N_i, int64_t N_p,
                                      const double *alpha, int K, int32_t *top_idx,
                                      float *top_val, void *stream)
{
    const char *fn = "mtip_match_topk_static";
    MTIP_REQUIRE(top_idx || N_i == 0, MTIP_ERR_ARG, "mtip_match_topk_static: null top_idx");
    int rc = mtip::check_match(fn, precision, slices, ld_s, slice_scale, slice_sumsq, N_o, images,
                               ld_d, image_scale, N_i, N_p, alpha, K);
    if (rc != MTIP_OK || N_i == 0) return rc;
    const int tile = mtip::gemm_tile_rows(precision, N_o, N_i);
    const size_t need = mtip::partial_bytes(N_o, N_i, K, tile);
    MTIP_REQUIRE(need <= (size_t)MTIP_STATIC_SCRATCH_BYTES, MTIP_ERR_ARG,
                 std::string(fn) + ": N_i * ceil(N_o / " + std::to_string(tile) + ") * K * 8 = " +
                     std::to_string(need) + " bytes of partial keys exceed the static scratch of " +
                     std::to_string((size_t)MTIP_STATIC_SCRATCH_BYTES) +
                     " bytes; match fewer images per call");
    uint64_t *part = nullptr;
    rc = scratch_on_current_device(&part);
    if (rc != MTIP_OK) return rc;
    hipStream_t st = (hipStream_t)stream;
    rc = mtip::partial_impl(fn, precision, slices, ld_s, slice_scale, slice_sumsq, N_o, images,
                            ld_d, image_scale, N_i, N_p, alpha, K, tile, part, st);
    if (rc != MTIP_OK) return rc;
    return mtip::merge_impl(part, N_o, N_i, K, tile, top_idx, top_val, st);
}
