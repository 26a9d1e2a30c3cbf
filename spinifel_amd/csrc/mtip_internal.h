// Internal helpers shared by the HIP sources of libmtip (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/mtip.h"

namespace mtip {

// Thread-local last error message (returned by mtip_last_error()).
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define MTIP_HIP_CHECK(expr)                                                     \
    do {                                                                         \
        hipError_t _e = (expr);                                                  \
        if (_e != hipSuccess)                                                    \
            return ::mtip::fail(MTIP_ERR_HIP, std::string(#expr) + ": " +       \
                                                  hipGetErrorString(_e));        \
    } while (0)

#define MTIP_REQUIRE(cond, code, msg)                                            \
    do {                                                                         \
        if (!(cond)) return ::mtip::fail((code), (msg));                         \
    } while (0)

// Order-preserving map of a float to uint32 (ascending float -> ascending
// uint, NaN above +inf), used for packed (value, index) keys.
__device__ __forceinline__ uint32_t float_to_ordered(float f)
{
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ordered_to_float(uint32_t u)
{
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}
__device__ __forceinline__ uint64_t make_key(float v, uint32_t idx)
{
    return ((uint64_t)float_to_ordered(v) << 32) | (uint64_t)idx;
}

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Host halves of the top-K matching (match.hip), shared with the other
// matching entry points: argument checks, the GEMM tile height chosen for a
// call, the partial-key bytes of a tile height, the MFMA GEMM writing
// per-tile top-K keys, and their merge into the per-image top-K.
int check_match(const char *fn, int precision, const void *slices, int64_t ld_s,
                const float *slice_scale, const double *slice_sumsq, int64_t N_o,
                const void *images, int64_t ld_d, const float *image_scale, int64_t N_i,
                int64_t N_p, const double *alpha, int K);
int gemm_tile_rows(int precision, int64_t N_o, int64_t N_i);
size_t partial_bytes(int64_t N_o, int64_t N_i, int K, int tile);
int partial_impl(const char *fn, int precision, const void *slices, int64_t ld_s,
                 const float *slice_scale, const double *slice_sumsq, int64_t N_o,
                 const void *images, int64_t ld_d, const float *image_scale, int64_t N_i,
                 int64_t N_p, const double *alpha, int K, int tile, uint64_t *part,
                 hipStream_t st, float *edge_ws = nullptr);
int merge_impl(const uint64_t *part, int64_t N_o, int64_t N_i, int K, int tile, int32_t *top_idx,
               float *top_val, hipStream_t st);

// 64-bit warp shuffle via two 32-bit shuffles.
__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int mask)
{
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    lo = __shfl_xor(lo, mask, 64);
    hi = __shfl_xor(hi, mask, 64);
    return ((uint64_t)hi << 32) | lo;
}

// Insert key into a sorted (ascending) array of K keys, dropping the largest.
template <int K>
__device__ __forceinline__ void topk_insert(uint64_t (&t)[K], uint64_t key)
{
    if (key < t[K - 1]) {
        t[K - 1] = key;
#pragma unroll
        for (int i = K - 1; i > 0; --i) {
            uint64_t lo = t[i - 1] < t[i] ? t[i - 1] : t[i];
            uint64_t hi = t[i - 1] < t[i] ? t[i] : t[i - 1];
            t[i - 1] = lo;
            t[i] = hi;
        }
    }
}

// Merge two ascending K-arrays (K a power of two), keep the K smallest in a,
// ascending.  min(a[i], b[K-1-i]) is a bitonic sequence holding the K
// smallest of the union; a bitonic merge network then sorts it.  All
// indices are compile-time constants so the arrays stay in registers.
template <int K>
__device__ __forceinline__ void topk_merge(uint64_t (&a)[K], const uint64_t (&b)[K])
{
#pragma unroll
    for (int i = 0; i < K; ++i) {
        uint64_t v = b[K - 1 - i];
        a[i] = a[i] < v ? a[i] : v;
    }
#pragma unroll
    for (int half = K / 2; half >= 1; half /= 2) {
#pragma unroll
        for (int i = 0; i < K; ++i) {
            if ((i & half) == 0) {
                uint64_t lo = a[i] < a[i + half] ? a[i] : a[i + half];
                uint64_t hi = a[i] < a[i + half] ? a[i + half] : a[i];
                a[i] = lo;
                a[i + half] = hi;
            }
        }
    }
}

}  // namespace mtip
