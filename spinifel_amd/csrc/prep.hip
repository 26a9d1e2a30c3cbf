// Data preparation of spinifel's prep step on gfx950: clipping to the
// low-resolution region followed by 2^n x 2^n binning, in one pass
// (spinifel/prep.py:36-134): every output pixel reads its block of the
// clipped input once and reduces it in the reference's order.
//
// Modes (the three binning flavours of prep.py):
//   0 sum   bin2x2_sum    ((x00 + x10) + x01) + x11 per level   (:36-40)
//   1 mean  bin2x2_mean   the same / 4 per level                (:44-45)
//   2 index bin2x2_index  min over row pairs, then column pairs,
//                         then floor(/ 2) per level             (:49-53)
// x_ij = arr[2r + i, 2c + j]; the sums are in the input type, so float
// results match numpy's bit for bit.
//
// Clipping (:73-105): with 1 panel the centred (sa / 2^c) x (sb / 2^c)
// window; with 4 panels the bottom-right (panels 0, 2) or bottom-left
// (panels 1, 3) corner of that size; any other panel count gives zeros, as
// the reference returns its zero array.
#include <type_traits>

#include "mtip_internal.h"

namespace {

template <typename T>
__device__ __forceinline__ T floordiv2(T v)
{
    if constexpr (std::is_integral<T>::value) {
        return v >= 0 ? v / 2 : -((-v + 1) / 2);
    } else {
        return floor(v / (T)2);
    }
}

// Reduction of the 2^L x 2^L block at p (row stride rs), mode as above.
template <typename T, int MODE, int L>
__device__ __forceinline__ T block(const T *__restrict__ p, int64_t rs)
{
    if constexpr (L == 0) {
        return p[0];
    } else {
        constexpr int h = 1 << (L - 1);
        const T x00 = block<T, MODE, L - 1>(p, rs);
        const T x10 = block<T, MODE, L - 1>(p + h * rs, rs);
        const T x01 = block<T, MODE, L - 1>(p + h, rs);
        const T x11 = block<T, MODE, L - 1>(p + h * rs + h, rs);
        if constexpr (MODE == 0) {
            return ((x00 + x10) + x01) + x11;
        } else if constexpr (MODE == 1) {
            return (((x00 + x10) + x01) + x11) / (T)4;
        } else {
            const T a = x00 < x10 ? x00 : x10;      // rows first
            const T b = x01 < x11 ? x01 : x11;
            return floordiv2<T>(a < b ? a : b);
        }
    }
}

template <typename T, int MODE, int L>
__global__ __launch_bounds__(256) void k_clip_bin(const T *__restrict__ in, T *__restrict__ out,
                                                  int64_t n_lead, int panels, int sa, int sb,
                                                  int ca, int cb, int oa, int ob)
{
    // ca x cb: clipped panel; oa x ob: output panel (ca >> L, cb >> L)
    const int64_t n_out = n_lead * panels * (int64_t)oa * ob;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n_out; i += stride) {
        const int c = (int)(i % ob);
        const int64_t t = i / ob;
        const int r = (int)(t % oa);
        const int64_t t2 = t / oa;
        const int pnl = (int)(t2 % panels);
        const int64_t lead = t2 / panels;
        // clipped window origin in the panel
        int r0, c0;
        if (panels == 1) {
            r0 = sa / 2 - ca / 2;
            c0 = sb / 2 - cb / 2;
        } else if (panels == 4) {
            r0 = sa - ca;
            c0 = (pnl == 0 || pnl == 2) ? sb - cb : 0;
        } else {
            out[i] = (T)0;
            continue;
        }
        const T *p = in + ((lead * panels + pnl) * (int64_t)sa + r0 + ((int64_t)r << L)) * sb + c0 +
                     ((int64_t)c << L);
        out[i] = block<T, MODE, L>(p, sb);
    }
}

template <typename T, int MODE>
int launch(const void *in, void *out, int64_t n_lead, int panels, int sa, int sb, int ca, int cb,
           int n_bin, hipStream_t st)
{
    const int oa = ca >> n_bin, ob = cb >> n_bin;
    const int64_t n_out = n_lead * panels * (int64_t)oa * ob;
    if (n_out == 0) return MTIP_OK;
    int64_t nb = mtip::ceil_div(n_out, 256);
    const unsigned grid = (unsigned)(nb < 65536 ? nb : 65536);
#define MTIP_CLIP_BIN(LV)                                                                        \
    hipLaunchKernelGGL((k_clip_bin<T, MODE, LV>), dim3(grid), dim3(256), 0, st, (const T *)in,   \
                       (T *)out, n_lead, panels, sa, sb, ca, cb, oa, ob)
    switch (n_bin) {
    case 0: MTIP_CLIP_BIN(0); break;
    case 1: MTIP_CLIP_BIN(1); break;
    case 2: MTIP_CLIP_BIN(2); break;
    case 3: MTIP_CLIP_BIN(3); break;
    case 4: MTIP_CLIP_BIN(4); break;
    default: return mtip::fail(MTIP_ERR_UNSUPPORTED, "mtip_clip_bin: n_bin must be in [0, 4]");
    }
#undef MTIP_CLIP_BIN
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}

}  // namespace

extern "C" int mtip_clip_bin(const void *in, int dtype, int64_t n_lead, int panels, int sa,
                             int sb, int n_clip, int n_bin, int mode, void *out, void *stream)
{
    MTIP_REQUIRE(n_lead >= 0 && panels >= 1 && sa >= 1 && sb >= 1, MTIP_ERR_ARG,
                 "mtip_clip_bin: bad shape");
    MTIP_REQUIRE(n_clip >= 0 && n_clip <= 8 && n_bin >= 0, MTIP_ERR_ARG,
                 "mtip_clip_bin: bad n_clip / n_bin");
    MTIP_REQUIRE(mode >= 0 && mode <= 2, MTIP_ERR_ARG, "mtip_clip_bin: mode must be 0, 1 or 2");
    // numpy's mean of an integer array is floating point: mean takes float data only
    MTIP_REQUIRE(mode != 1 || dtype <= 1, MTIP_ERR_ARG, "mtip_clip_bin: mean needs float data");
    const int f = 1 << n_clip;
    MTIP_REQUIRE(sa % f == 0 && sb % f == 0, MTIP_ERR_ARG,
                 "mtip_clip_bin: panel edges must be multiples of 2^n_clip");
    // the centred window of a 1-panel detector is sa / 2^c about sa / 2
    MTIP_REQUIRE(panels != 1 || ((sa / f) % 2 == 0 && (sb / f) % 2 == 0) || n_clip == 0,
                 MTIP_ERR_ARG, "mtip_clip_bin: centred window must have even edges");
    const int ca = sa / f, cb = sb / f;
    MTIP_REQUIRE(n_bin <= 4 && ca % (1 << n_bin) == 0 && cb % (1 << n_bin) == 0, MTIP_ERR_ARG,
                 "mtip_clip_bin: clipped edges must be multiples of 2^n_bin (n_bin <= 4)");
    if (n_lead == 0) return MTIP_OK;
    MTIP_REQUIRE(in && out, MTIP_ERR_ARG, "mtip_clip_bin: null pointer");
    hipStream_t st = (hipStream_t)stream;
#define MTIP_CB_MODE(T)                                                                        \
    switch (mode) {                                                                            \
    case 0: return launch<T, 0>(in, out, n_lead, panels, sa, sb, ca, cb, n_bin, st);           \
    case 1: return launch<T, 1>(in, out, n_lead, panels, sa, sb, ca, cb, n_bin, st);           \
    default: return launch<T, 2>(in, out, n_lead, panels, sa, sb, ca, cb, n_bin, st);          \
    }
    switch (dtype) {
    case 0: MTIP_CB_MODE(float)
    case 1: MTIP_CB_MODE(double)
    case 2: MTIP_CB_MODE(int32_t)
    case 3: MTIP_CB_MODE(int64_t)
    default: return mtip::fail(MTIP_ERR_ARG, "mtip_clip_bin: dtype must be 0 (f32), 1 (f64), 2 (i32), 3 (i64)");
    }
#undef MTIP_CB_MODE
}
