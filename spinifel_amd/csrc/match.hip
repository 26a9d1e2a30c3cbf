// Slice-to-image distances on gfx950: MFMA cross-term GEMM with a fused
// distance epilogue and per-image top-K, top-K merge, FP32/FP64 rescoring,
// and the small row utilities around them.
//
// Distance expansion (per image i, orientation n, scaled slices alpha s_n):
//   ||alpha s_n - d_i||^2 = alpha^2 ||s_n||^2 - 2 alpha s_n . d_i + ||d_i||^2
// The last term is constant per image, so ranking uses
//   v(n, i) = alpha^2 ||s_n||^2 - 2 alpha (s_n . d_i)
// and the rescoring kernel recomputes the full distance for candidates.
//
// GEMM tiling: a 256-thread workgroup owns a 128 (orientations) x 128
// (images) tile; waves form a 2 x 2 grid of 64 x 64 sub-tiles, each a 2 x 2
// set of 32 x 32 MFMA blocks.  K is staged 128 bytes per row at a time
// (32 fp32 / 64 bf16 / 128 e4m3) through LDS with a register prefetch of the
// next stage.  Operand fragments: the lower half-wave (lanes 0-31) feeds the
// first half of each MFMA's k-range and the upper half the second; A and B
// use the same k assignment, so every pixel is paired with itself whatever
// order the hardware sums in.
#include <math.h>

#include <atomic>

#include "mtip_internal.h"

namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int intx8 __attribute__((ext_vector_type(8)));

constexpr int BM = 128, BN = 128, BKB = 128;
constexpr int LDS_STRIDE = BKB + 16;      // bytes per staged row (conflict-free b128 reads)
constexpr uint64_t KEY_NONE = ~0ull;

enum { OUT_TOPK = 0, OUT_FULL = 1 };

template <int PREC> struct Prec;
template <> struct Prec<MTIP_FP32> { static constexpr int bytes = 4; };
template <> struct Prec<MTIP_BF16> { static constexpr int bytes = 2; };
template <> struct Prec<MTIP_FP8> { static constexpr int bytes = 1; };

__device__ __forceinline__ uint4 load16(const uint8_t *p) { return *(const uint4 *)p; }

// One MFMA step on a 32 x 32 block.  `off` is the byte offset of this
// step inside the staged 128-byte row.
template <int PREC>
__device__ __forceinline__ void mfma_stage(const uint8_t *sA_row0, const uint8_t *sA_row1,
                                           const uint8_t *sB_row0, const uint8_t *sB_row1,
                                           int half, floatx16 (&acc)[2][2])
{
    if constexpr (PREC == MTIP_FP32) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int off = g * 32 + half * 16;
            floatx4 a0 = *(const floatx4 *)(sA_row0 + off);
            floatx4 a1 = *(const floatx4 *)(sA_row1 + off);
            floatx4 b0 = *(const floatx4 *)(sB_row0 + off);
            floatx4 b1 = *(const floatx4 *)(sB_row1 + off);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0[t], b0[t], acc[0][0], 0, 0, 0);
                acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0[t], b1[t], acc[0][1], 0, 0, 0);
                acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1[t], b0[t], acc[1][0], 0, 0, 0);
                acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1[t], b1[t], acc[1][1], 0, 0, 0);
            }
        }
    } else if constexpr (PREC == MTIP_BF16) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int off = s * 32 + half * 16;
            bf16x8 a0 = *(const bf16x8 *)(sA_row0 + off);
            bf16x8 a1 = *(const bf16x8 *)(sA_row1 + off);
            bf16x8 b0 = *(const bf16x8 *)(sB_row0 + off);
            bf16x8 b1 = *(const bf16x8 *)(sB_row1 + off);
            acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc[1][1], 0, 0, 0);
        }
    } else {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int off = s * 64 + half * 32;
            intx8 a0 = *(const intx8 *)(sA_row0 + off);
            intx8 a1 = *(const intx8 *)(sA_row1 + off);
            intx8 b0 = *(const intx8 *)(sB_row0 + off);
            intx8 b1 = *(const intx8 *)(sB_row1 + off);
            // format 0 = e4m3 for both operands; E8M0 block scales 127 = x1
            acc[0][0] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a0, b0, acc[0][0], 0, 0, 0, 127, 0, 127);
            acc[0][1] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a0, b1, acc[0][1], 0, 0, 0, 127, 0, 127);
            acc[1][0] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a1, b0, acc[1][0], 0, 0, 0, 127, 0, 127);
            acc[1][1] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a1, b1, acc[1][1], 0, 0, 0, 127, 0, 127);
        }
    }
}

constexpr int BM2 = 256, BN2 = 256;
constexpr int kXcds = 8;

__device__ __forceinline__ void tile_of(int bid, int nblocks, int nta, int ntb, int &ta, int &tb)
{
    // workgroup bid runs on XCD bid % 8 as its (bid / 8)-th; XCD x runs q + 1
    // workgroups when x < rem, else q (nblocks = 8 q + rem), and takes the
    // contiguous tile range that starts at x q + min(x, rem): a bijection of
    // [0, nblocks) for any nblocks
    const int q = nblocks / kXcds, rem = nblocks % kXcds;
    const int xcd = bid % kXcds, local = bid / kXcds;
    const int t = xcd * q + min(xcd, rem) + local;
    constexpr int GM = 8;
    const int in_group = GM * ntb;
    const int group = t / in_group;
    const int first = group * GM;
    const int gm = min(nta - first, GM);
    const int r = t % in_group;
    ta = first + r % gm;
    tb = r / gm;
}

// A: rows = orientations (slices), B: rows = images.  kbytes = N_p * bytes.
template <int PREC, int K, int OUT>
__global__ __launch_bounds__(256) void k_gemm(const uint8_t *__restrict__ A, int64_t lda,
                                              int64_t NA, const uint8_t *__restrict__ B,
                                              int64_t ldb, int64_t NB, int64_t kbytes,
                                              const float *__restrict__ scaleA,
                                              const float *__restrict__ scaleB,
                                              const double *__restrict__ sumsqA,
                                              const double *__restrict__ alpha, int n_tiles_a,
                                              uint64_t *__restrict__ part, float *__restrict__ full,
                                              int64_t ldf)
{
    __shared__ __attribute__((aligned(16))) uint8_t smem[(BM + BN) * LDS_STRIDE];
    __shared__ float s_rowc[BM];     // TOPK: alpha^2 ||s||^2 ; FULL: ||a||^2
    __shared__ float s_colc[BN];     // FULL: ||b||^2
    __shared__ float s_sa[BM], s_sb[BN];
    uint8_t *sA = smem;
    uint8_t *sB = smem + BM * LDS_STRIDE;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int half = lane >> 5, l32 = lane & 31;
    const int ta = blockIdx.x % n_tiles_a, tb = blockIdx.x / n_tiles_a;
    const int64_t a0 = (int64_t)ta * BM, b0 = (int64_t)tb * BN;

    if (tid < BM) {
        int64_t r = a0 + tid;
        float sc = 1.0f;
        if (PREC == MTIP_FP8 && r < NA) sc = scaleA[r];
        s_sa[tid] = sc;
        if (OUT == OUT_TOPK) {
            double al = *alpha;
            s_rowc[tid] = r < NA ? (float)(al * al * sumsqA[r]) : 0.0f;
        }
    } else {
        int64_t r = b0 + (tid - BM);
        float sc = 1.0f;
        if (PREC == MTIP_FP8 && r < NB) sc = scaleB[r];
        s_sb[tid - BM] = sc;
    }

    floatx16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;

    // staging: 1024 16-byte chunks per operand tile, 4 per thread
    uint4 ra[4], rb[4];
    auto load_stage = [&](int64_t kb) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int c = tid + 256 * i;
            int r = c >> 3, col = (c & 7) * 16;
            int64_t gk = kb + col;
            int64_t ga = a0 + r, gb = b0 + r;
            ra[i] = (ga < NA && gk < kbytes) ? load16(A + ga * lda + gk) : make_uint4(0, 0, 0, 0);
            rb[i] = (gb < NB && gk < kbytes) ? load16(B + gb * ldb + gk) : make_uint4(0, 0, 0, 0);
        }
    };
    double nrm = 0.0;    // FULL: FP64 squared norm of this thread's staged row

    load_stage(0);
    for (int64_t kb = 0; kb < kbytes; kb += BKB) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int c = tid + 256 * i;
            int r = c >> 3, col = (c & 7) * 16;
            *(uint4 *)(sA + r * LDS_STRIDE + col) = ra[i];
            *(uint4 *)(sB + r * LDS_STRIDE + col) = rb[i];
        }
        __syncthreads();
        if (kb + BKB < kbytes) load_stage(kb + BKB);
        if (OUT == OUT_FULL) {
            const float *row = (const float *)((tid < BM ? sA + tid * LDS_STRIDE
                                                         : sB + (tid - BM) * LDS_STRIDE));
#pragma unroll
            for (int e = 0; e < BKB / 4; ++e) nrm += (double)row[e] * (double)row[e];
        }
        const uint8_t *ar0 = sA + (wm * 64 + l32) * LDS_STRIDE;
        const uint8_t *ar1 = sA + (wm * 64 + 32 + l32) * LDS_STRIDE;
        const uint8_t *br0 = sB + (wn * 64 + l32) * LDS_STRIDE;
        const uint8_t *br1 = sB + (wn * 64 + 32 + l32) * LDS_STRIDE;
        mfma_stage<PREC>(ar0, ar1, br0, br1, half, acc);
    }

    if constexpr (OUT == OUT_FULL) {
        __syncthreads();
        if (tid < BM) s_rowc[tid] = (float)nrm;
        else s_colc[tid - BM] = (float)nrm;
        __syncthreads();
        const bool vec_ok = (ldf % 4) == 0;
#pragma unroll
        for (int bn = 0; bn < 2; ++bn) {
            const int j = wn * 64 + bn * 32 + l32;
            const int64_t gi = b0 + j;
            if (gi >= NB) continue;
            const float nb = s_colc[j];
#pragma unroll
            for (int bm = 0; bm < 2; ++bm) {
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int m = wm * 64 + bm * 32 + 8 * g + 4 * half;
                    const int64_t go = a0 + m;
                    float o[4];
#pragma unroll
                    for (int t = 0; t < 4; ++t)
                        o[t] = fmaxf(s_rowc[m + t] + nb - 2.0f * acc[bm][bn][4 * g + t], 0.0f);
                    float *dst = full + gi * ldf + go;
                    if (vec_ok && go + 3 < NA) {
                        *(floatx4 *)dst = floatx4{o[0], o[1], o[2], o[3]};
                    } else {
#pragma unroll
                        for (int t = 0; t < 4; ++t)
                            if (go + t < NA) dst[t] = o[t];
                    }
                }
            }
        }
    } else {
        const float two_a = (float)(2.0 * (*alpha));
        uint64_t top[2][K];
#pragma unroll
        for (int bn = 0; bn < 2; ++bn) {
#pragma unroll
            for (int r = 0; r < K; ++r) top[bn][r] = KEY_NONE;
            const int j = wn * 64 + bn * 32 + l32;
            const float sb = s_sb[j];
#pragma unroll
            for (int bm = 0; bm < 2; ++bm) {
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int m = wm * 64 + bm * 32 + 8 * (v >> 2) + 4 * half + (v & 3);
                    const int64_t go = a0 + m;
                    if (go < NA) {
                        float c = acc[bm][bn][v];
                        if (PREC == MTIP_FP8) c *= s_sa[m] * sb;
                        float val = fmaf(-two_a, c, s_rowc[m]);
                        mtip::topk_insert<K>(top[bn], mtip::make_key(val, (uint32_t)go));
                    }
                }
            }
            uint64_t other[K];
#pragma unroll
            for (int r = 0; r < K; ++r) other[r] = mtip::shfl_xor_u64(top[bn][r], 32);
            mtip::topk_merge<K>(top[bn], other);
        }
        // cross-wave merge over the two 64-row halves (wm = 0, 1)
        __syncthreads();
        uint64_t *s_top = (uint64_t *)smem;    // [2 wn][64 cols][K]
        if (wm == 1 && half == 0) {
#pragma unroll
            for (int bn = 0; bn < 2; ++bn)
#pragma unroll
                for (int r = 0; r < K; ++r)
                    s_top[(wn * 64 + bn * 32 + l32) * K + r] = top[bn][r];
        }
        __syncthreads();
        if (wm == 0 && half == 0) {
#pragma unroll
            for (int bn = 0; bn < 2; ++bn) {
                const int j = wn * 64 + bn * 32 + l32;
                uint64_t other[K];
#pragma unroll
                for (int r = 0; r < K; ++r) other[r] = s_top[j * K + r];
                mtip::topk_merge<K>(top[bn], other);
                const int64_t gi = b0 + j;
                if (gi < NB) {
                    uint64_t *dst = part + (gi * n_tiles_a + ta) * K;
#pragma unroll
                    for (int r = 0; r < K; ++r) dst[r] = top[bn][r];
                }
            }
        }
    }
}

// ---------------------------------------------------------------- big tiles
//
// 256 (orientations) x 256 (images) per 512-thread workgroup for large
// problems: 8 waves in a 2 (M) x 4 (N) grid of 128 x 64 sub-tiles, each a
// 4 x 2 set of 32 x 32 MFMA blocks.  Two LDS stages (double buffer, one
// barrier per K step) with the next-but-one stage prefetched into registers.
// Twice the flops per staged byte of the 128 x 128 kernel, which is bound by
// L2 -> LDS traffic at bf16 / e4m3 rates.  Tile order: workgroup ids are
// dealt to the 8 XCDs round robin, so each XCD gets a contiguous range of
// tiles, walked 8 tile-rows at a time, and the workgroups running together
// on one XCD share A and B panels in its L2.
template <int PREC, int K, bool EDGE>
__global__ __launch_bounds__(512) void k_gemm_big(const uint8_t *__restrict__ A, int64_t lda,
                                                  int64_t NA, const uint8_t *__restrict__ B,
                                                  int64_t ldb, int64_t NB, int64_t kbytes,
                                                  const float *__restrict__ scaleA,
                                                  const float *__restrict__ scaleB,
                                                  const double *__restrict__ sumsqA,
                                                  const double *__restrict__ alpha, int nta,
                                                  int ntb, uint64_t *__restrict__ part, int nwa,
                                                  int nwb, int splits, float *__restrict__ spart,
                                                  int slot)
{
    __shared__ __attribute__((aligned(16))) uint8_t smem[2 * (BM2 + BN2) * LDS_STRIDE];
    __shared__ float s_rowc[BM2];
    __shared__ float s_sa[BM2], s_sb[BN2];
    constexpr int STAGE = (BM2 + BN2) * LDS_STRIDE;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int half = lane >> 5, l32 = lane & 31;
    // EDGE = false: the nwa x nwb whole tiles (nwa = nta, nwb = ntb unless
    // the thin edge tiles go to the EDGE launch; the key layout keeps nta
    // tile rows).  EDGE = true: each thin edge tile (last row of orientation
    // tiles, then last column of image tiles) as `splits` workgroups over
    // consecutive parts of K that write their raw sums to spart
    int ta, tb, kp = 0;
    if constexpr (!EDGE) {
        tile_of(blockIdx.x, nwa * nwb, nwa, nwb, ta, tb);
    } else {
        const int t = (int)blockIdx.x / splits, n_row = nwa < nta ? ntb : 0;
        kp = (int)blockIdx.x - t * splits;
        ta = t < n_row ? nta - 1 : t - n_row;
        tb = t < n_row ? t : ntb - 1;
    }
    const int64_t a0 = (int64_t)ta * BM2, b0 = (int64_t)tb * BN2;

    if (tid < BM2) {
        const int64_t r = a0 + tid;
        float sc = 1.0f;
        if (PREC == MTIP_FP8 && r < NA) sc = scaleA[r];
        s_sa[tid] = sc;
        const double al = *alpha;
        s_rowc[tid] = r < NA ? (float)(al * al * sumsqA[r]) : 0.0f;
    } else {
        const int64_t r = b0 + (tid - BM2);
        float sc = 1.0f;
        if (PREC == MTIP_FP8 && r < NB) sc = scaleB[r];
        s_sb[tid - BM2] = sc;
    }

    floatx16 acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;

    // staging: 2048 16-byte chunks per operand tile, 4 per thread per operand,
    // as buffer loads over the tile's own rows: rows past the matrix and
    // columns past kbytes are out of range and read as zero, so the loads
    // need no branches (the host keeps 256 rows x ld within 2^31 bytes)
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    const int rowsA = (int)min((int64_t)BM2, NA - a0), rowsB = (int)min((int64_t)BN2, NB - b0);
    const __amdgpu_buffer_rsrc_t rsA = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(A + a0 * lda), (short)0, (int)((int64_t)rowsA * lda), 0x00020000);
    const __amdgpu_buffer_rsrc_t rsB = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(B + b0 * ldb), (short)0, (int)((int64_t)rowsB * ldb), 0x00020000);
    const int colb = (tid & 7) * 16;
    const int vA = (tid >> 3) * (int)lda + colb, vB = (tid >> 3) * (int)ldb + colb;
    // stages [0, nstage) from byte kb0; loads at or past kend read zeros
    int nstage = (int)((kbytes + BKB - 1) / BKB), kb0 = 0;
    int64_t kend = kbytes;
    if constexpr (EDGE) {
        const int s0 = (int)((int64_t)kp * nstage / splits);
        const int s1 = (int)((int64_t)(kp + 1) * nstage / splits);
        kb0 = s0 * BKB;
        kend = min(kbytes, (int64_t)s1 * BKB);
        nstage = s1 - s0;
    }
    u32x4 ra[4], rb[4];
    auto load_stage = [&](int64_t kb) __attribute__((always_inline)) {
        const bool in_k = kb + colb < kend;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int oa = in_k ? vA + 64 * i * (int)lda + (int)kb : 0x7fffffff;
            const int ob = in_k ? vB + 64 * i * (int)ldb + (int)kb : 0x7fffffff;
            ra[i] = __builtin_amdgcn_raw_buffer_load_b128(rsA, oa, 0, 0);
            rb[i] = __builtin_amdgcn_raw_buffer_load_b128(rsB, ob, 0, 0);
        }
    };
    auto store_stage = [&](int buf) __attribute__((always_inline)) {
        uint8_t *sA = smem + buf * STAGE;
        uint8_t *sB = sA + BM2 * LDS_STRIDE;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int c = tid + 512 * i;
            const int r = c >> 3, col = (c & 7) * 16;
            *(u32x4 *)(sA + r * LDS_STRIDE + col) = ra[i];
            *(u32x4 *)(sB + r * LDS_STRIDE + col) = rb[i];
        }
    };

    load_stage(kb0);
    store_stage(0);
    if (nstage > 1) load_stage(kb0 + BKB);
    __syncthreads();
    for (int st = 0; st < nstage; ++st) {
        // stage st + 1 (fetched during stage st - 1) goes to the other buffer,
        // last read in stage st - 1, and stage st + 2 is fetched: both overlap
        // this stage's MFMAs; the barrier at the end publishes stage st + 1
        // (past the last stage these write zeros to the idle buffer and fetch
        // nothing: the loop body stays one straight-line block)
        store_stage((st + 1) & 1);
        load_stage((int64_t)kb0 + (int64_t)(st + 2) * BKB);
        const uint8_t *sA = smem + (st & 1) * STAGE;
        const uint8_t *sB = sA + BM2 * LDS_STRIDE;
        const uint8_t *ar[4], *br[2];
#pragma unroll
        for (int i = 0; i < 4; ++i) ar[i] = sA + (wm * 128 + i * 32 + l32) * LDS_STRIDE;
#pragma unroll
        for (int j = 0; j < 2; ++j) br[j] = sB + (wn * 64 + j * 32 + l32) * LDS_STRIDE;
        if constexpr (PREC == MTIP_FP32) {
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int off = g * 32 + half * 16;
                floatx4 av[4], bv[2];
#pragma unroll
                for (int i = 0; i < 4; ++i) av[i] = *(const floatx4 *)(ar[i] + off);
#pragma unroll
                for (int j = 0; j < 2; ++j) bv[j] = *(const floatx4 *)(br[j] + off);
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int j = 0; j < 2; ++j)
                            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[i][t], bv[j][t], acc[i][j], 0, 0, 0);
            }
        } else if constexpr (PREC == MTIP_BF16) {
            // fragments of sub-step s4 + 1 are read while sub-step s4's MFMAs run
            bf16x8 av[2][4], bv[2][2];
            auto frag = [&](int s4, int b) __attribute__((always_inline)) {
                const int off = s4 * 32 + half * 16;
#pragma unroll
                for (int i = 0; i < 4; ++i) av[b][i] = *(const bf16x8 *)(ar[i] + off);
#pragma unroll
                for (int j = 0; j < 2; ++j) bv[b][j] = *(const bf16x8 *)(br[j] + off);
            };
            frag(0, 0);
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                if (s4 + 1 < 4) frag(s4 + 1, (s4 + 1) & 1);
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[s4 & 1][i], bv[s4 & 1][j], acc[i][j], 0, 0, 0);
            }
            // issue order: sub-step 0's fragments, then each MFMA followed by
            // one fragment read of the next sub-step (18 in all), with the 8
            // LDS stores and 8 loads of the staging spread over the stage
            __builtin_amdgcn_sched_group_barrier(0x100, 6, 0);
#define MTIP_SG(DSR, DSW, VM)                                                  \
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                         \
    if (DSR) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);                \
    if (DSW) __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);                \
    if (VM) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
#define MTIP_SG4(R)  MTIP_SG(R, 1, 0) MTIP_SG(R, 0, 1) MTIP_SG(R, 0, 0) MTIP_SG(R, 0, 0)
            MTIP_SG4(1) MTIP_SG4(1) MTIP_SG4(1) MTIP_SG4(1)
            MTIP_SG(1, 1, 0) MTIP_SG(1, 0, 1) MTIP_SG(0, 0, 0) MTIP_SG(0, 0, 0)
            MTIP_SG4(0) MTIP_SG4(0) MTIP_SG4(0)
#undef MTIP_SG4
#undef MTIP_SG
        } else {
#pragma unroll
            for (int s2 = 0; s2 < 2; ++s2) {
                const int off = s2 * 64 + half * 32;
                intx8 bv[2];
#pragma unroll
                for (int j = 0; j < 2; ++j) bv[j] = *(const intx8 *)(br[j] + off);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const intx8 av = *(const intx8 *)(ar[i] + off);
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv[j], acc[i][j], 0, 0, 0, 127, 0, 127);
                }
            }
            // issue order: sub-step 0's fragment reads, then each MFMA followed
            // by reads of sub-step 1 (12 over the first 8 MFMAs), with the
            // staging's 8 LDS stores and then its 8 loads spread over the stage
            __builtin_amdgcn_sched_group_barrier(0x100, 12, 0);
#define MTIP_SG(DSR, DSW, VM)                                                  \
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                         \
    if (DSR) __builtin_amdgcn_sched_group_barrier(0x100, DSR, 0);              \
    if (DSW) __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);                \
    if (VM) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
            MTIP_SG(2, 1, 0) MTIP_SG(1, 1, 0) MTIP_SG(2, 1, 0) MTIP_SG(1, 1, 0)
            MTIP_SG(2, 1, 0) MTIP_SG(1, 1, 0) MTIP_SG(2, 1, 0) MTIP_SG(1, 1, 0)
            MTIP_SG(0, 0, 1) MTIP_SG(0, 0, 1) MTIP_SG(0, 0, 1) MTIP_SG(0, 0, 1)
            MTIP_SG(0, 0, 1) MTIP_SG(0, 0, 1) MTIP_SG(0, 0, 1) MTIP_SG(0, 0, 1)
#undef MTIP_SG
        }
        __syncthreads();
    }

    if constexpr (EDGE) {
        // this part of K's raw sums over the tile's valid rows x columns
        float *dst = spart + (int64_t)blockIdx.x * slot;
#pragma unroll
        for (int bn = 0; bn < 2; ++bn) {
            const int n = wn * 64 + bn * 32 + l32;
#pragma unroll
            for (int bm = 0; bm < 4; ++bm)
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int m = wm * 128 + bm * 32 + 8 * (v >> 2) + 4 * half + (v & 3);
                    if (m < rowsA && n < rowsB) dst[m * rowsB + n] = acc[bm][bn][v];
                }
        }
        return;
    }
    // epilogue: per image column the K smallest v over the tile's 256 rows
    const float two_a = 2.0f * (float)(*alpha);
    uint64_t top[2][K];
#pragma unroll
    for (int bn = 0; bn < 2; ++bn) {
#pragma unroll
        for (int r = 0; r < K; ++r) top[bn][r] = KEY_NONE;
        const int j = wn * 64 + bn * 32 + l32;
        const float sb = s_sb[j];
#pragma unroll
        for (int bm = 0; bm < 4; ++bm) {
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int m = wm * 128 + bm * 32 + 8 * (v >> 2) + 4 * half + (v & 3);
                const int64_t go = a0 + m;
                if (go < NA) {
                    float c = acc[bm][bn][v];
                    if (PREC == MTIP_FP8) c *= s_sa[m] * sb;
                    const float val = fmaf(-two_a, c, s_rowc[m]);
                    mtip::topk_insert<K>(top[bn], mtip::make_key(val, (uint32_t)go));
                }
            }
        }
        uint64_t other[K];
#pragma unroll
        for (int r = 0; r < K; ++r) other[r] = mtip::shfl_xor_u64(top[bn][r], 32);
        mtip::topk_merge<K>(top[bn], other);
    }
    uint64_t *s_top = (uint64_t *)smem;    // [4 wn][64 cols][K] (after the final barrier)
    if (wm == 1 && half == 0) {
#pragma unroll
        for (int bn = 0; bn < 2; ++bn)
#pragma unroll
            for (int r = 0; r < K; ++r) s_top[(wn * 64 + bn * 32 + l32) * K + r] = top[bn][r];
    }
    __syncthreads();
    if (wm == 0 && half == 0) {
#pragma unroll
        for (int bn = 0; bn < 2; ++bn) {
            const int j = wn * 64 + bn * 32 + l32;
            uint64_t other[K];
#pragma unroll
            for (int r = 0; r < K; ++r) other[r] = s_top[j * K + r];
            mtip::topk_merge<K>(top[bn], other);
            const int64_t gi = b0 + j;
            if (gi < NB) {
                uint64_t *dst = part + (gi * nta + ta) * K;
#pragma unroll
                for (int r = 0; r < K; ++r) dst[r] = top[bn][r];
            }
        }
    }
}

// Thin edge tiles computed as `splits` parts of K: per element the parts
// summed in order, then the whole-tile epilogue's value and the per-column
// top-K.  One workgroup per tile; a column's rows are shared by tpc
// consecutive lanes (256 / columns, a power of two up to 16), whose lists
// are then merged by xor shuffles.
template <int PREC, int K>
__global__ __launch_bounds__(256) void k_edge_fixup(const float *__restrict__ spart, int splits,
                                                    int slot, int nta, int ntb, int nwa,
                                                    int64_t NA, int64_t NB,
                                                    const float *__restrict__ scaleA,
                                                    const float *__restrict__ scaleB,
                                                    const double *__restrict__ sumsqA,
                                                    const double *__restrict__ alpha,
                                                    uint64_t *__restrict__ part)
{
    const int t = blockIdx.x, n_row = nwa < nta ? ntb : 0;
    const int ta = t < n_row ? nta - 1 : t - n_row, tb = t < n_row ? t : ntb - 1;
    const int64_t a0 = (int64_t)ta * BM2, b0 = (int64_t)tb * BN2;
    const int rA = (int)min((int64_t)BM2, NA - a0), rB = (int)min((int64_t)BN2, NB - b0);
    int tpc = 1;
    while (tpc < 16 && tpc * 2 * rB <= 256) tpc *= 2;
    const int n = threadIdx.x / tpc, g = threadIdx.x % tpc;
    const double al = *alpha;
    const float two_a = 2.0f * (float)al;
    uint64_t top[K];
#pragma unroll
    for (int r = 0; r < K; ++r) top[r] = KEY_NONE;
    if (n < rB) {
        const float sb = PREC == MTIP_FP8 ? scaleB[b0 + n] : 1.0f;
        const float *src = spart + (int64_t)t * splits * slot + n;
        for (int m = g; m < rA; m += tpc) {
            float c = src[m * rB];
            for (int p = 1; p < splits; ++p) c += src[(int64_t)p * slot + m * rB];
            if (PREC == MTIP_FP8) c *= scaleA[a0 + m] * sb;
            const float rowc = (float)(al * al * sumsqA[a0 + m]);
            mtip::topk_insert<K>(top, mtip::make_key(fmaf(-two_a, c, rowc), (uint32_t)(a0 + m)));
        }
    }
    for (int d = 1; d < tpc; d <<= 1) {
        uint64_t other[K];
#pragma unroll
        for (int r = 0; r < K; ++r) other[r] = mtip::shfl_xor_u64(top[r], d);
        mtip::topk_merge<K>(top, other);
    }
    if (n < rB && g == 0) {
        uint64_t *dst = part + ((b0 + n) * nta + ta) * K;
#pragma unroll
        for (int r = 0; r < K; ++r) dst[r] = top[r];
    }
}

// One wave per row: K smallest keys of `n_keys` packed keys (KEYS = true)
// or of `n_keys` floats (KEYS = false, key = (value, column)).
template <int K, bool KEYS>
__global__ __launch_bounds__(256) void k_topk_rows(const uint64_t *__restrict__ keys,
                                                   const float *__restrict__ vals,
                                                   int64_t n_rows, int64_t n_keys, int64_t ld,
                                                   int K_out, int32_t *__restrict__ idx,
                                                   float *__restrict__ val)
{
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n_rows) return;
    uint64_t t[K];
#pragma unroll
    for (int r = 0; r < K; ++r) t[r] = KEY_NONE;
    for (int64_t e = lane; e < n_keys; e += 64) {
        uint64_t k = KEYS ? keys[row * ld + e] : mtip::make_key(vals[row * ld + e], (uint32_t)e);
        mtip::topk_insert<K>(t, k);
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        uint64_t o[K];
#pragma unroll
        for (int r = 0; r < K; ++r) o[r] = mtip::shfl_xor_u64(t[r], m);
        mtip::topk_merge<K>(t, o);
    }
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < K; ++r) {
            if (r >= K_out) break;
            bool none = t[r] == KEY_NONE;
            if (idx) idx[row * K_out + r] = none ? -1 : (int32_t)(uint32_t)t[r];
            if (val) val[row * K_out + r] = none ? __builtin_nanf("") : mtip::ordered_to_float((uint32_t)(t[r] >> 32));
        }
    }
}

__device__ __forceinline__ double wave_sum_f64(double v)
{
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}
__device__ __forceinline__ float wave_sum_f32(float v)
{
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// Block per image: full distance to each of K candidates in FP32 and FP64.
template <int K>
__global__ __launch_bounds__(256) void k_rescore(const float *__restrict__ S, int64_t lds,
                                                 int64_t N_o, const float *__restrict__ D, int64_t ldd,
                                                 int64_t N_p, const double *__restrict__ alpha,
                                                 int alpha_per_image,
                                                 const int32_t *__restrict__ cand,
                                                 int32_t *__restrict__ best_idx,
                                                 double *__restrict__ best_d64,
                                                 float *__restrict__ d32, double *__restrict__ d64,
                                                 const int32_t *__restrict__ order)
{
    __shared__ double s64[4][K];
    __shared__ float s32[4][K];
    // images in the caller's order (e.g. grouped by best candidate, so that
    // neighbouring workgroups read the same candidate slices); workgroups go
    // round-robin over the 8 XCDs, so each XCD takes a contiguous run of the
    // order and the images resident on it share its L2
    int64_t i = blockIdx.x;
    if (order) {
        const int64_t nb = gridDim.x, n8 = nb / 8, b = blockIdx.x;
        i = (int64_t)order[b < 8 * n8 ? (b % 8) * n8 + b / 8 : b];
    }
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double al = alpha_per_image ? alpha[i] : alpha[0];
    const float alf = (float)al;
    int32_t c[K];
#pragma unroll
    for (int r = 0; r < K; ++r) {
        int32_t v = cand[i * K + r];
        c[r] = (v >= 0 && v < N_o) ? v : -1;
    }
    double a64[K];
    float a32[K];
#pragma unroll
    for (int r = 0; r < K; ++r) { a64[r] = 0.0; a32[r] = 0.0f; }
    const float *d = D + i * ldd;
    const bool vec = (N_p % 4 == 0) && (ldd % 4 == 0) && (lds % 4 == 0);
    if (vec) {
        for (int64_t p = 4 * tid; p < N_p; p += 1024) {
            floatx4 dv = *(const floatx4 *)(d + p);
#pragma unroll
            for (int r = 0; r < K; ++r) {
                if (c[r] < 0) continue;
                floatx4 sv = *(const floatx4 *)(S + (int64_t)c[r] * lds + p);
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    double e = al * (double)sv[t] - (double)dv[t];
                    a64[r] += e * e;
                    float ef = fmaf(alf, sv[t], -dv[t]);
                    a32[r] = fmaf(ef, ef, a32[r]);
                }
            }
        }
    } else {
        for (int64_t p = tid; p < N_p; p += 256) {
            float dv = d[p];
#pragma unroll
            for (int r = 0; r < K; ++r) {
                if (c[r] < 0) continue;
                float sv = S[(int64_t)c[r] * lds + p];
                double e = al * (double)sv - (double)dv;
                a64[r] += e * e;
                float ef = fmaf(alf, sv, -dv);
                a32[r] = fmaf(ef, ef, a32[r]);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < K; ++r) {
        double v64 = wave_sum_f64(a64[r]);
        float v32 = wave_sum_f32(a32[r]);
        if (lane == 0) { s64[wave][r] = v64; s32[wave][r] = v32; }
    }
    __syncthreads();
    if (tid == 0) {
        int32_t bi = -1;
        double bd = 0.0;
#pragma unroll
        for (int r = 0; r < K; ++r) {
            double t64 = s64[0][r] + s64[1][r] + s64[2][r] + s64[3][r];
            float t32 = (s32[0][r] + s32[1][r]) + (s32[2][r] + s32[3][r]);
            if (c[r] < 0) { t64 = __builtin_nan(""); t32 = __builtin_nanf(""); }
            if (d64) d64[i * K + r] = t64;
            if (d32) d32[i * K + r] = t32;
            if (c[r] >= 0 && (bi < 0 || t64 < bd || (t64 == bd && c[r] < bi))) {
                bi = c[r];
                bd = t64;
            }
        }
        if (best_idx) best_idx[i] = bi;
        if (best_d64) best_d64[i] = bi < 0 ? __builtin_nan("") : bd;
    }
}

__global__ __launch_bounds__(256) void k_row_stats(const float *__restrict__ in, int64_t cols,
                                                   int64_t ld, double *__restrict__ rs,
                                                   double *__restrict__ rq)
{
    __shared__ double sa[4], sb[4];
    const int64_t r = blockIdx.x;
    double s = 0.0, q = 0.0;
    for (int64_t c = threadIdx.x; c < cols; c += 256) {
        double v = in[r * ld + c];
        s += v;
        q += v * v;
    }
    s = wave_sum_f64(s);
    q = wave_sum_f64(q);
    if ((threadIdx.x & 63) == 0) { sa[threadIdx.x >> 6] = s; sb[threadIdx.x >> 6] = q; }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (rs) rs[r] = (sa[0] + sa[1]) + (sa[2] + sa[3]);
        if (rq) rq[r] = (sb[0] + sb[1]) + (sb[2] + sb[3]);
    }
}

__global__ __launch_bounds__(256) void k_scaling_ratio(const double *__restrict__ is,
                                                       const double *__restrict__ iq, int64_t ni,
                                                       const double *__restrict__ ss,
                                                       const double *__restrict__ sq, int64_t ns,
                                                       int64_t n_pix, double *__restrict__ alpha)
{
    __shared__ double red[4][4];
    double a = 0, b = 0, c = 0, d = 0;
    for (int64_t k = threadIdx.x; k < ni; k += 256) { a += is[k]; b += iq[k]; }
    for (int64_t k = threadIdx.x; k < ns; k += 256) { c += ss[k]; d += sq[k]; }
    a = wave_sum_f64(a);
    b = wave_sum_f64(b);
    c = wave_sum_f64(c);
    d = wave_sum_f64(d);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[w][0] = a; red[w][1] = b; red[w][2] = c; red[w][3] = d; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[4];
        for (int k = 0; k < 4; ++k) t[k] = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
        double Ni = (double)ni * (double)n_pix, Ns = (double)ns * (double)n_pix;
        double mi = t[0] / Ni, ms = t[2] / Ns;
        double vi = fmax(t[1] / Ni - mi * mi, 0.0), vs = fmax(t[3] / Ns - ms * ms, 0.0);
        *alpha = sqrt(vi) / sqrt(vs);
    }
}

// alpha per image: for each segment [off[g], off[g+1]) of images (one request),
// std(images of the segment) / std(all slices), written to every image of it.
__global__ __launch_bounds__(256) void k_segment_ratio(const double *__restrict__ is,
                                                       const double *__restrict__ iq,
                                                       const int64_t *__restrict__ off,
                                                       const double *__restrict__ ss,
                                                       const double *__restrict__ sq, int64_t ns,
                                                       int64_t n_pix, double *__restrict__ alpha)
{
    __shared__ double red[4][4];
    const int64_t g = blockIdx.x;
    const int64_t a0 = off[g], a1 = off[g + 1];
    double a = 0, b = 0, c = 0, d = 0;
    // unrolled so that the loads of several iterations are in flight at
    // once (the sums keep their order)
#pragma unroll 8
    for (int64_t k = a0 + threadIdx.x; k < a1; k += 256) { a += is[k]; b += iq[k]; }
#pragma unroll 8
    for (int64_t k = threadIdx.x; k < ns; k += 256) { c += ss[k]; d += sq[k]; }
    a = wave_sum_f64(a);
    b = wave_sum_f64(b);
    c = wave_sum_f64(c);
    d = wave_sum_f64(d);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[w][0] = a; red[w][1] = b; red[w][2] = c; red[w][3] = d; }
    __syncthreads();
    double t[4];
    for (int k = 0; k < 4; ++k) t[k] = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
    const double Ni = (double)(a1 - a0) * (double)n_pix, Ns = (double)ns * (double)n_pix;
    const double mi = t[0] / Ni, ms = t[2] / Ns;
    const double vi = fmax(t[1] / Ni - mi * mi, 0.0), vs = fmax(t[3] / Ns - ms * ms, 0.0);
    const double r = sqrt(vi) / sqrt(vs);
    for (int64_t k = a0 + threadIdx.x; k < a1; k += 256) alpha[k] = r;
}

__device__ __forceinline__ uint16_t bf16_rne(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ float block_max_f32(float m, float *red)
{
    for (int s = 32; s >= 1; s >>= 1) m = fmaxf(m, __shfl_xor(m, s, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    return fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

// Block per row: r = m * x - c (m from a device scalar or its inverse, c a
// per-column centre), cast to FMT, FP64 sum of squares of the values the
// GEMM will see (dequantised), FP8 per-row scale absmax / 448.  Columns in
// [cols, ld_out) are written as zero.
template <int FMT>
__global__ __launch_bounds__(256) void k_residual(const float *__restrict__ in, int64_t cols,
                                                  int64_t ld_in, const double *__restrict__ mult,
                                                  int invert, const float *__restrict__ center,
                                                  void *__restrict__ out, int64_t ld_out,
                                                  float *__restrict__ scales,
                                                  double *__restrict__ sumsq, int vec)
{
    __shared__ float redf[4];
    __shared__ double redd[4];
    const int64_t r = blockIdx.x;
    const float *src = in + r * ld_in;
    float m = 1.0f;
    if (mult) {
        const double mv = (invert & 2) ? mult[r] : mult[0];
        m = (invert & 1) ? (float)(1.0 / mv) : (float)mv;
    }
    // Thread t owns columns 4t + 1024 k + (0..3).  Each trip handles four such
    // groups, k = 4j .. 4j + 3, with all their loads issued before any use, so
    // a block keeps 16 KB of its row in flight; sums keep the column order.
    constexpr int kGroups = 4;
    const int64_t cols4 = (cols / 4) * 4;
    // x = m src - centre over columns c4 .. c4 + 3 (zero at and past cols)
    auto load4 = [&](int64_t c4, float (&x)[4]) __attribute__((always_inline)) {
        if (vec && c4 + 4 <= cols4) {
            const float4 v = *reinterpret_cast<const float4 *>(src + c4);
            const float vv[4] = {v.x, v.y, v.z, v.w};
            if (center) {
                const float4 cc = *reinterpret_cast<const float4 *>(center + c4);
                const float cv[4] = {cc.x, cc.y, cc.z, cc.w};
#pragma unroll
                for (int t = 0; t < 4; ++t) x[t] = fmaf(m, vv[t], -cv[t]);
            } else {
#pragma unroll
                for (int t = 0; t < 4; ++t) x[t] = m * vv[t];
            }
        } else {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int64_t c = c4 + t;
                x[t] = c < cols ? (center ? fmaf(m, src[c], -center[c]) : m * src[c]) : 0.0f;
            }
        }
    };
    const int64_t c_first = 4 * (int64_t)threadIdx.x;
    float scale = 1.0f, inv = 1.0f;
    if (FMT == MTIP_FP8) {
        float a = 0.0f;
        for (int64_t c0 = c_first; c0 < cols; c0 += 1024 * kGroups) {
            float x[kGroups][4];
#pragma unroll
            for (int u = 0; u < kGroups; ++u) load4(c0 + 1024 * u, x[u]);
#pragma unroll
            for (int u = 0; u < kGroups; ++u)
#pragma unroll
                for (int t = 0; t < 4; ++t) a = fmaxf(a, fabsf(x[u][t]));
        }
        a = block_max_f32(a, redf);
        scale = (a > 0.0f && isfinite(a)) ? a / 448.0f : 1.0f;
        inv = 1.0f / scale;
        if (threadIdx.x == 0 && scales) scales[r] = scale;
    }
    double ss = 0.0;
    for (int64_t c0 = c_first; c0 < ld_out; c0 += 1024 * kGroups) {
        float x[kGroups][4];
#pragma unroll
        for (int u = 0; u < kGroups; ++u) load4(c0 + 1024 * u, x[u]);
#pragma unroll
        for (int u = 0; u < kGroups; ++u) {
            const int64_t c4 = c0 + 1024 * u;
            if (c4 >= ld_out) break;
            const bool full = c4 + 4 <= ld_out;
            if (FMT == MTIP_FP8) {
                float v[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) v[t] = fminf(fmaxf(x[u][t] * inv, -448.0f), 448.0f);
                int pk = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], 0, false);
                pk = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], pk, true);
                const float dqv[4] = {__builtin_amdgcn_cvt_f32_fp8(pk, 0), __builtin_amdgcn_cvt_f32_fp8(pk, 1),
                                      __builtin_amdgcn_cvt_f32_fp8(pk, 2), __builtin_amdgcn_cvt_f32_fp8(pk, 3)};
                uint8_t *d = (uint8_t *)out + r * ld_out + c4;
                if (vec && full) *reinterpret_cast<uint32_t *>(d) = (uint32_t)pk;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    if (c4 + t < ld_out) {
                        if (!(vec && full)) d[t] = (uint8_t)((uint32_t)pk >> (8 * t));
                        const double dq = (double)dqv[t] * (double)scale;
                        ss += dq * dq;
                    }
                }
            } else if (FMT == MTIP_BF16) {
                uint16_t b[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) b[t] = bf16_rne(x[u][t]);
                uint16_t *d = (uint16_t *)out + r * ld_out + c4;
                if (vec && full) {
                    uint2 pk;
                    pk.x = (uint32_t)b[0] | ((uint32_t)b[1] << 16);
                    pk.y = (uint32_t)b[2] | ((uint32_t)b[3] << 16);
                    *reinterpret_cast<uint2 *>(d) = pk;
                }
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    if (c4 + t < ld_out) {
                        if (!(vec && full)) d[t] = b[t];
                        const double dq = bf16_to_f32(b[t]);
                        ss += dq * dq;
                    }
                }
            } else {
                float *d = (float *)out + r * ld_out + c4;
                if (vec && full)
                    *reinterpret_cast<float4 *>(d) = make_float4(x[u][0], x[u][1], x[u][2], x[u][3]);
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    if (c4 + t < ld_out) {
                        if (!(vec && full)) d[t] = x[u][t];
                        ss += (double)x[u][t] * (double)x[u][t];
                    }
                }
            }
        }
    }
    if (sumsq) {
        ss = wave_sum_f64(ss);
        __syncthreads();
        if ((threadIdx.x & 63) == 0) redd[threadIdx.x >> 6] = ss;
        __syncthreads();
        if (threadIdx.x == 0) sumsq[r] = (redd[0] + redd[1]) + (redd[2] + redd[3]);
    }
}

// Column sums (FP64) of a float32 (rows, cols) matrix over kColGroups row
// groups: block of 256 = 64 columns x 4 row phases, grid (column blocks, row
// groups); partial sums land in work[group][col] and are added in a fixed
// order by k_col_finish, so the result does not depend on scheduling.
constexpr int kColGroups = 16;

__global__ __launch_bounds__(256) void k_col_partial(const float *__restrict__ in, int64_t rows,
                                                     int64_t cols, int64_t ld,
                                                     double *__restrict__ work)
{
    __shared__ double part[4][64];
    const int cx = threadIdx.x & 63, ph = threadIdx.x >> 6;
    const int64_t c = (int64_t)blockIdx.x * 64 + cx;
    const int64_t per = (rows + kColGroups - 1) / kColGroups;
    const int64_t r0 = (int64_t)blockIdx.y * per, r1 = min(rows, r0 + per);
    double s0 = 0.0, s1 = 0.0;
    if (c < cols) {
        int64_t r = r0 + ph;
        for (; r + 4 < r1; r += 8) {
            s0 += in[r * ld + c];
            s1 += in[(r + 4) * ld + c];
        }
        for (; r < r1; r += 4) s0 += in[r * ld + c];
    }
    part[ph][cx] = s0 + s1;
    __syncthreads();
    if (ph == 0 && c < cols)
        work[(int64_t)blockIdx.y * cols + c] = (part[0][cx] + part[1][cx]) + (part[2][cx] + part[3][cx]);
}

__global__ __launch_bounds__(256) void k_col_finish(const double *__restrict__ work, int64_t rows,
                                                    int64_t cols, float *__restrict__ mean)
{
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= cols) return;
    double t = 0.0;
    for (int g = 0; g < kColGroups; ++g) t += work[(int64_t)g * cols + c];
    mean[c] = (float)(t / (double)rows);
}


// Per-precision tile override (0 = rule below), set by mtip_set_gemm_tile.
std::atomic<int> g_tile_override[3] = {{0}, {0}, {0}};

// Orientation rows per top-K partial tile: 256 (big kernel) for large BF16
// and FP8 problems, else 128, unless overridden.  The partial buffer and the
// merge use the same value.
int tile_rows(int precision, int64_t N_o, int64_t N_i)
{
    if (precision >= 0 && precision < 3) {
        const int o = g_tile_override[precision].load(std::memory_order_relaxed);
        if (o) return o;
    }
    // the 256-row tile once it has about a round of tiles to spread over the
    // CUs (measured at 16,384 px, BF16: 250 or 500 orientations x 58,642
    // images, 230 / 460 tiles, 0.58 / 1.17 ms against 0.84 / 1.68 on the 128
    // tile; 128 tiles of 256 tie or lose)
    const int64_t tiles = mtip::ceil_div(N_o, BM2) * mtip::ceil_div(N_i, BN2);
    return (precision != MTIP_FP32 && tiles >= 192) ? BM2 : BM;
}

// Thin edge tiles of the 256-row kernel: the last row of orientation tiles
// or the last column of image tiles when it holds at most kEdgeRows valid
// rows.  A thin tile costs as long as a whole one (its time is the K loop's
// staging, not its MFMAs), so at 10,000 x 10,000 the 79 thin tiles of 1,600
// add a seventh, quarter-full round of tiles on 256 CUs.  Split along K, they
// run as one short extra launch after the 1,521 whole tiles.
constexpr int kEdgeRows = 64;

struct EdgePlan {
    int nta, ntb, nwa, nwb, n_thin, splits, slot;
};

EdgePlan edge_plan(int64_t N_o, int64_t N_i)
{
    EdgePlan p;
    p.nta = (int)mtip::ceil_div(N_o, BM2);
    p.ntb = (int)mtip::ceil_div(N_i, BN2);
    const int ea = (int)(N_o - (int64_t)(p.nta - 1) * BM2);
    const int eb = (int)(N_i - (int64_t)(p.ntb - 1) * BN2);
    const bool thin_a = p.nta > 1 && ea <= kEdgeRows, thin_b = p.ntb > 1 && eb <= kEdgeRows;
    p.nwa = thin_a ? p.nta - 1 : p.nta;
    p.nwb = thin_b ? p.ntb - 1 : p.ntb;
    p.n_thin = (thin_a ? p.ntb : 0) + (thin_b ? p.nwa : 0);
    // about one round of workgroups over the CUs, each a part of K
    p.splits = p.n_thin ? std::max(1, std::min(8, 256 / p.n_thin)) : 1;
    p.slot = std::max(thin_a ? ea * BN2 : 0, thin_b ? BM2 * eb : 0);
    return p;
}

// Bytes of FP32 partial sums the split edge tiles need (0: none to split).
size_t edge_bytes(int64_t N_o, int64_t N_i)
{
    if (N_o <= 0 || N_i <= 0) return 0;
    const EdgePlan p = edge_plan(N_o, N_i);
    return (size_t)p.n_thin * p.splits * p.slot * sizeof(float);
}

template <int PREC, int K>
void launch_gemm_topk(const void *S, int64_t lds_b, int64_t N_o, const void *D, int64_t ldd_b,
                      int64_t N_i, int64_t kbytes, const float *ssc, const float *dsc,
                      const double *s2, const double *alpha, int tile, uint64_t *part,
                      float *edge_ws, hipStream_t st)
{
    if (tile == BM2) {
        EdgePlan p = edge_plan(N_o, N_i);
        if (!edge_ws || !p.n_thin) {
            p.nwa = p.nta;
            p.nwb = p.ntb;
            p.n_thin = 0;
        }
        hipLaunchKernelGGL((k_gemm_big<PREC, K, false>), dim3((unsigned)(p.nwa * p.nwb)),
                           dim3(512), 0, st, (const uint8_t *)S, lds_b, N_o, (const uint8_t *)D,
                           ldd_b, N_i, kbytes, ssc, dsc, s2, alpha, p.nta, p.ntb, part, p.nwa,
                           p.nwb, 1, nullptr, 0);
        if (p.n_thin) {
            hipLaunchKernelGGL((k_gemm_big<PREC, K, true>), dim3((unsigned)(p.n_thin * p.splits)),
                               dim3(512), 0, st, (const uint8_t *)S, lds_b, N_o,
                               (const uint8_t *)D, ldd_b, N_i, kbytes, ssc, dsc, s2, alpha, p.nta,
                               p.ntb, part, p.nwa, p.nwb, p.splits, edge_ws, p.slot);
            hipLaunchKernelGGL((k_edge_fixup<PREC, K>), dim3((unsigned)p.n_thin), dim3(256), 0, st,
                               edge_ws, p.splits, p.slot, p.nta, p.ntb, p.nwa, N_o, N_i, ssc, dsc,
                               s2, alpha, part);
        }
    } else {
        const int nta = (int)mtip::ceil_div(N_o, BM), ntb = (int)mtip::ceil_div(N_i, BN);
        hipLaunchKernelGGL((k_gemm<PREC, K, OUT_TOPK>), dim3((unsigned)(nta * ntb)), dim3(256), 0,
                           st, (const uint8_t *)S, lds_b, N_o, (const uint8_t *)D, ldd_b, N_i,
                           kbytes, ssc, dsc, s2, alpha, nta, part, nullptr, (int64_t)0);
    }
}

// Indices of the K smallest entries of each row for any K: one workgroup per
// row, output in chunks of up to kSelChunk keys.  Keys are (order-preserving
// float bits << 32 | column), so they are unique and ties go to the lowest
// column.  Per chunk, the rank-r key T among the keys above the previous
// chunk's last key L is found by a radix select (eight passes of 8-bit
// digits, each a histogram of the row in LDS); the keys in (L, T] are then
// gathered, bitonic-sorted in LDS and written.  Columns past the row give -1.
constexpr int kSelChunk = 2048;

__global__ __launch_bounds__(256) void k_select_rows(const float *__restrict__ vals, int64_t n_rows,
                                                     int64_t n, int64_t ld, int64_t K,
                                                     int32_t *__restrict__ idx)
{
    __shared__ uint32_t hist[256];
    __shared__ uint64_t buf[kSelChunk];
    __shared__ uint64_t s_prefix;
    __shared__ int64_t s_rank;
    __shared__ uint32_t s_count;
    const int tid = threadIdx.x;
    const int64_t row = blockIdx.x;
    if (row >= n_rows) return;
    const float *v = vals + row * ld;
    int32_t *out = idx + row * K;
    bool have_lo = false;
    uint64_t lo = 0;                    // last key written (exclusive lower bound)
    for (int64_t done = 0; done < K;) {
        // keys above lo
        if (tid == 0) s_count = 0;
        __syncthreads();
        uint32_t c = 0;
        for (int64_t e = tid; e < n; e += 256) {
            const uint64_t k = mtip::make_key(v[e], (uint32_t)e);
            c += (!have_lo || k > lo) ? 1u : 0u;
        }
        atomicAdd(&s_count, c);
        __syncthreads();
        const int64_t avail = s_count;
        const int64_t want = min((int64_t)kSelChunk, K - done);
        const int64_t take = min(want, avail);
        if (take == 0) {
            for (int64_t r = done + tid; r < K; r += 256) out[r] = -1;
            return;
        }
        // radix select of the take-th smallest key above lo
        if (tid == 0) { s_prefix = 0; s_rank = take; }
        for (int pass = 0; pass < 8; ++pass) {
            const int shift = 56 - 8 * pass;
            hist[tid] = 0;
            __syncthreads();
            const uint64_t prefix = s_prefix;
            for (int64_t e = tid; e < n; e += 256) {
                const uint64_t k = mtip::make_key(v[e], (uint32_t)e);
                if ((have_lo && k <= lo) || (pass > 0 && (k >> (shift + 8)) != prefix)) continue;
                atomicAdd(&hist[(k >> shift) & 255], 1u);
            }
            __syncthreads();
            if (tid == 0) {
                int64_t r = s_rank;
                int d = 0;
                for (; d < 255 && (int64_t)hist[d] < r; ++d) r -= hist[d];
                s_prefix = (prefix << 8) | (uint64_t)d;
                s_rank = r;
            }
            __syncthreads();
        }
        const uint64_t hi = s_prefix;   // the take-th key: keys in (lo, hi] are this chunk
        if (tid == 0) s_count = 0;
        for (int i = tid; i < kSelChunk; i += 256) buf[i] = KEY_NONE;
        __syncthreads();
        for (int64_t e = tid; e < n; e += 256) {
            const uint64_t k = mtip::make_key(v[e], (uint32_t)e);
            if ((!have_lo || k > lo) && k <= hi) buf[atomicAdd(&s_count, 1u)] = k;
        }
        __syncthreads();
        int len = 1;
        while (len < take) len <<= 1;
        for (int size = 2; size <= len; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int i = tid; i < len; i += 256) {
                    const int j = i ^ stride;
                    if (j > i) {
                        const bool up = (i & size) == 0;
                        const uint64_t a = buf[i], b = buf[j];
                        if ((a > b) == up) { buf[i] = b; buf[j] = a; }
                    }
                }
                __syncthreads();
            }
        }
        for (int64_t r = tid; r < take; r += 256) out[done + r] = (int32_t)(uint32_t)buf[r];
        __syncthreads();
        lo = hi;
        have_lo = true;
        done += take;
    }
}

template <int K>
void launch_topk_rows(const uint64_t *keys, const float *vals, int64_t rows, int64_t n, int64_t ld,
                      int kout, int32_t *idx, float *val, hipStream_t st)
{
    dim3 g((unsigned)mtip::ceil_div(rows, 4));
    if (keys)
        hipLaunchKernelGGL((k_topk_rows<K, true>), g, dim3(256), 0, st, keys, vals, rows, n, ld,
                           kout, idx, val);
    else
        hipLaunchKernelGGL((k_topk_rows<K, false>), g, dim3(256), 0, st, keys, vals, rows, n, ld,
                           kout, idx, val);
}

int prec_bytes(int p) { return p == MTIP_FP32 ? 4 : p == MTIP_BF16 ? 2 : p == MTIP_FP8 ? 1 : 0; }

}  // namespace

extern "C" size_t mtip_match_workspace_size(int64_t N_o, int64_t N_i, int K, int precision)
{
    (void)precision;
    if (N_o <= 0 || N_i <= 0 || K <= 0) return 0;
    int64_t nta = mtip::ceil_div(N_o, BM);
    return (size_t)(N_i * nta * K) * sizeof(uint64_t);
}

namespace {

int check_match_args(const char *fn, int precision, const void *slices, int64_t ld_s,
                     const float *slice_scale, const double *slice_sumsq, int64_t N_o,
                     const void *images, int64_t ld_d, const float *image_scale, int64_t N_i,
                     int64_t N_p, const double *alpha, int K)
{
    const std::string f(fn);
    const int eb = prec_bytes(precision);
    MTIP_REQUIRE(eb > 0, MTIP_ERR_ARG, f + ": unknown precision");
    MTIP_REQUIRE(K == 1 || K == 2 || K == 4 || K == 8, MTIP_ERR_ARG, f + ": K must be 1, 2, 4 or 8");
    MTIP_REQUIRE(N_o >= 0 && N_i >= 0 && N_p >= 0, MTIP_ERR_ARG, f + ": negative size");
    MTIP_REQUIRE(N_o <= 0x7fffffff, MTIP_ERR_ARG, f + ": N_o too large");
    if (N_i == 0) return MTIP_OK;
    MTIP_REQUIRE(N_o > 0, MTIP_ERR_ARG, f + ": no orientations");
    MTIP_REQUIRE(N_p > 0 && N_p % 16 == 0, MTIP_ERR_ARG, f + ": N_p must be a positive multiple of 16");
    MTIP_REQUIRE(ld_s >= N_p && ld_d >= N_p && (ld_s * eb) % 16 == 0 && (ld_d * eb) % 16 == 0,
                 MTIP_ERR_ARG, f + ": leading dimensions must be >= N_p and 16-byte multiples");
    MTIP_REQUIRE(slices && images && slice_sumsq && alpha, MTIP_ERR_ARG, f + ": null pointer");
    MTIP_REQUIRE(precision != MTIP_FP8 || (slice_scale && image_scale), MTIP_ERR_ARG,
                 f + ": FP8 needs per-row scales");
    MTIP_REQUIRE(((uintptr_t)slices % 16) == 0 && ((uintptr_t)images % 16) == 0, MTIP_ERR_ARG,
                 f + ": operands must be 16-byte aligned");
    const int64_t nta = mtip::ceil_div(N_o, BM), ntb = mtip::ceil_div(N_i, BN);
    MTIP_REQUIRE(nta * ntb < 0x7fffffff, MTIP_ERR_ARG, f + ": grid too large");
    return MTIP_OK;
}

}  // namespace

extern "C" int mtip_set_gemm_tile(int precision, int rows)
{
    MTIP_REQUIRE(precision == MTIP_FP32 || precision == MTIP_BF16 || precision == MTIP_FP8,
                 MTIP_ERR_ARG, "mtip_set_gemm_tile: unknown precision");
    MTIP_REQUIRE(rows == 0 || rows == BM || rows == BM2, MTIP_ERR_ARG,
                 "mtip_set_gemm_tile: rows must be 0, 128 or 256");
    g_tile_override[precision].store(rows, std::memory_order_relaxed);
    return MTIP_OK;
}

extern "C" int mtip_gemm_tile(int precision, int64_t N_o, int64_t N_i)
{
    MTIP_REQUIRE(precision == MTIP_FP32 || precision == MTIP_BF16 || precision == MTIP_FP8,
                 MTIP_ERR_ARG, "mtip_gemm_tile: unknown precision");
    return tile_rows(precision, N_o, N_i);
}

namespace mtip {

// Tile rows the caller asked for (0: the rule and override of tile_rows()).
int resolve_tile(int precision, int64_t N_o, int64_t N_i, int tile)
{
    return tile == 0 ? tile_rows(precision, N_o, N_i) : tile;
}

// Bytes of top-K partial keys for a given tile height.
size_t partial_bytes(int64_t N_o, int64_t N_i, int K, int tile)
{
    return (size_t)(N_i * mtip::ceil_div(N_o, tile) * K) * sizeof(uint64_t);
}

int partial_impl(const char *fn, int precision, const void *slices, int64_t ld_s,
                 const float *slice_scale, const double *slice_sumsq, int64_t N_o,
                 const void *images, int64_t ld_d, const float *image_scale, int64_t N_i,
                 int64_t N_p, const double *alpha, int K, int tile, uint64_t *part,
                 hipStream_t st, float *edge_ws)
{
    const std::string f(fn);
    const int eb = prec_bytes(precision);
    // the 256-row kernel addresses a tile's rows with 32-bit buffer offsets
    MTIP_REQUIRE(tile != BM2 || (int64_t)BM2 * (ld_s > ld_d ? ld_s : ld_d) * eb < ((int64_t)1 << 31),
                 MTIP_ERR_UNSUPPORTED,
                 f + ": rows longer than 8 MiB need the 128-row tile "
                     "(mtip_set_gemm_tile(precision, 128))");
    const int64_t kbytes = N_p * eb;
#define MTIP_GEMM_CASE(P)                                                                        \
    switch (K) {                                                                                 \
    case 1: launch_gemm_topk<P, 1>(slices, ld_s * eb, N_o, images, ld_d * eb, N_i, kbytes,     \
                                   slice_scale, image_scale, slice_sumsq, alpha, tile, part, edge_ws, st); \
            break;                                                                               \
    case 2: launch_gemm_topk<P, 2>(slices, ld_s * eb, N_o, images, ld_d * eb, N_i, kbytes,     \
                                   slice_scale, image_scale, slice_sumsq, alpha, tile, part, edge_ws, st); \
            break;                                                                               \
    case 4: launch_gemm_topk<P, 4>(slices, ld_s * eb, N_o, images, ld_d * eb, N_i, kbytes,     \
                                   slice_scale, image_scale, slice_sumsq, alpha, tile, part, edge_ws, st); \
            break;                                                                               \
    default: launch_gemm_topk<P, 8>(slices, ld_s * eb, N_o, images, ld_d * eb, N_i, kbytes,    \
                                    slice_scale, image_scale, slice_sumsq, alpha, tile, part, edge_ws, st); \
             break;                                                                              \
    }
    if (precision == MTIP_FP32) { MTIP_GEMM_CASE(MTIP_FP32) }
    else if (precision == MTIP_BF16) { MTIP_GEMM_CASE(MTIP_BF16) }
    else { MTIP_GEMM_CASE(MTIP_FP8) }
#undef MTIP_GEMM_CASE
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}

int merge_impl(const uint64_t *part, int64_t N_o, int64_t N_i, int K, int tile, int32_t *top_idx,
               float *top_val, hipStream_t st)
{
    const int64_t nkeys = mtip::ceil_div(N_o, tile) * K;
    switch (K) {
    case 1: launch_topk_rows<1>(part, nullptr, N_i, nkeys, nkeys, K, top_idx, top_val, st); break;
    case 2: launch_topk_rows<2>(part, nullptr, N_i, nkeys, nkeys, K, top_idx, top_val, st); break;
    case 4: launch_topk_rows<4>(part, nullptr, N_i, nkeys, nkeys, K, top_idx, top_val, st); break;
    default: launch_topk_rows<8>(part, nullptr, N_i, nkeys, nkeys, K, top_idx, top_val, st); break;
    }
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}

int check_tile(const char *fn, int tile)
{
    MTIP_REQUIRE(tile == 0 || tile == BM || tile == BM2, MTIP_ERR_ARG,
                 std::string(fn) + ": tile_rows must be 0, 128 or 256");
    return MTIP_OK;
}

int gemm_tile_rows(int precision, int64_t N_o, int64_t N_i)
{
    return tile_rows(precision, N_o, N_i);
}

int check_match(const char *fn, int precision, const void *slices, int64_t ld_s,
                const float *slice_scale, const double *slice_sumsq, int64_t N_o,
                const void *images, int64_t ld_d, const float *image_scale, int64_t N_i,
                int64_t N_p, const double *alpha, int K)
{
    return check_match_args(fn, precision, slices, ld_s, slice_scale, slice_sumsq, N_o, images,
                            ld_d, image_scale, N_i, N_p, alpha, K);
}

}  // namespace mtip

using mtip::check_tile;
using mtip::merge_impl;
using mtip::partial_bytes;
using mtip::partial_impl;
using mtip::resolve_tile;

namespace {

int match_partial(const char *fn, int precision, const void *slices, int64_t ld_s,
                  const float *slice_scale, const double *slice_sumsq, int64_t N_o,
                  const void *images, int64_t ld_d, const float *image_scale, int64_t N_i,
                  int64_t N_p, const double *alpha, int K, int tile_rows_arg, int flags,
                  void *workspace, size_t workspace_bytes, void *stream)
{
    const std::string f(fn);
    int rc = check_match_args(fn, precision, slices, ld_s, slice_scale, slice_sumsq, N_o, images,
                              ld_d, image_scale, N_i, N_p, alpha, K);
    if (rc == MTIP_OK) rc = check_tile(fn, tile_rows_arg);
    if (rc != MTIP_OK) return rc;
    MTIP_REQUIRE((flags & ~MTIP_MATCH_SPLIT_EDGES) == 0, MTIP_ERR_ARG, f + ": unknown flags");
    if (N_i == 0) return MTIP_OK;
    const int tile = resolve_tile(precision, N_o, N_i, tile_rows_arg);
    size_t need = partial_bytes(N_o, N_i, K, tile);
    float *edge_ws = nullptr;
    if ((flags & MTIP_MATCH_SPLIT_EDGES) && tile == BM2) {
        // split only where the partial sums fit the size mtip_match_workspace_size
        // gives, so that the choice depends on the sizes alone
        const size_t keys = (need + 255) & ~(size_t)255, extra = edge_bytes(N_o, N_i);
        if (extra && keys + extra <= mtip_match_workspace_size(N_o, N_i, K, precision)) {
            need = keys + extra;
            edge_ws = (float *)((char *)workspace + keys);
        }
    }
    MTIP_REQUIRE(workspace && workspace_bytes >= need, MTIP_ERR_NOMEM, f + ": workspace too small");
    return partial_impl(fn, precision, slices, ld_s, slice_scale, slice_sumsq, N_o, images, ld_d,
                        image_scale, N_i, N_p, alpha, K, tile, (uint64_t *)workspace,
                        (hipStream_t)stream, edge_ws);
}

}  // namespace

extern "C" int mtip_match_partial(int precision, const void *slices, int64_t ld_s,
                                  const float *slice_scale, const double *slice_sumsq, int64_t N_o,
                                  const void *images, int64_t ld_d, const float *image_scale,
                                  int64_t N_i, int64_t N_p, const double *alpha, int K,
                                  int tile_rows_arg, void *workspace, size_t workspace_bytes,
                                  void *stream)
{
    return match_partial("mtip_match_partial", precision, slices, ld_s, slice_scale, slice_sumsq,
                         N_o, images, ld_d, image_scale, N_i, N_p, alpha, K, tile_rows_arg, 0,
                         workspace, workspace_bytes, stream);
}

extern "C" int mtip_match_partial_ex(int precision, const void *slices, int64_t ld_s,
                                     const float *slice_scale, const double *slice_sumsq,
                                     int64_t N_o, const void *images, int64_t ld_d,
                                     const float *image_scale, int64_t N_i, int64_t N_p,
                                     const double *alpha, int K, int tile_rows_arg, int flags,
                                     void *workspace, size_t workspace_bytes, void *stream)
{
    return match_partial("mtip_match_partial_ex", precision, slices, ld_s, slice_scale,
                         slice_sumsq, N_o, images, ld_d, image_scale, N_i, N_p, alpha, K,
                         tile_rows_arg, flags, workspace, workspace_bytes, stream);
}

extern "C" int mtip_topk_merge(int precision, const void *workspace, int64_t N_o, int64_t N_i,
                               int K, int tile_rows_arg, int32_t *top_idx, float *top_val,
                               void *stream)
{
    MTIP_REQUIRE(precision == MTIP_FP32 || precision == MTIP_BF16 || precision == MTIP_FP8,
                 MTIP_ERR_ARG, "mtip_topk_merge: unknown precision");
    MTIP_REQUIRE(K == 1 || K == 2 || K == 4 || K == 8, MTIP_ERR_ARG, "mtip_topk_merge: K must be 1, 2, 4 or 8");
    MTIP_REQUIRE(N_o > 0 && N_i >= 0, MTIP_ERR_ARG, "mtip_topk_merge: bad size");
    int rc = check_tile("mtip_topk_merge", tile_rows_arg);
    if (rc != MTIP_OK) return rc;
    if (N_i == 0) return MTIP_OK;
    MTIP_REQUIRE(workspace && top_idx, MTIP_ERR_ARG, "mtip_topk_merge: null pointer");
    return merge_impl((const uint64_t *)workspace, N_o, N_i, K,
                      resolve_tile(precision, N_o, N_i, tile_rows_arg), top_idx, top_val,
                      (hipStream_t)stream);
}

extern "C" int mtip_match_topk(int precision, const void *slices, int64_t ld_s,
                               const float *slice_scale, const double *slice_sumsq, int64_t N_o,
                               const void *images, int64_t ld_d, const float *image_scale,
                               int64_t N_i, int64_t N_p, const double *alpha, int K,
                               void *workspace, size_t workspace_bytes, int32_t *top_idx,
                               float *top_val, void *stream)
{
    MTIP_REQUIRE(top_idx || N_i == 0, MTIP_ERR_ARG, "mtip_match_topk: null top_idx");
    int rc = check_match_args("mtip_match_topk", precision, slices, ld_s, slice_scale,
                              slice_sumsq, N_o, images, ld_d, image_scale, N_i, N_p, alpha, K);
    if (rc != MTIP_OK || N_i == 0) return rc;
    // one tile height for both halves, read once
    const int tile = tile_rows(precision, N_o, N_i);
    MTIP_REQUIRE(workspace && workspace_bytes >= partial_bytes(N_o, N_i, K, tile), MTIP_ERR_NOMEM,
                 "mtip_match_topk: workspace too small");
    hipStream_t st = (hipStream_t)stream;
    rc = partial_impl("mtip_match_topk", precision, slices, ld_s, slice_scale, slice_sumsq, N_o,
                      images, ld_d, image_scale, N_i, N_p, alpha, K, tile, (uint64_t *)workspace,
                      st);
    if (rc != MTIP_OK) return rc;
    return merge_impl((const uint64_t *)workspace, N_o, N_i, K, tile, top_idx, top_val, st);
}

extern "C" int mtip_rescore(const float *slices, int64_t ld_s, int64_t N_o, const float *images, int64_t ld_d,
                            int64_t N_i, int64_t N_p, const double *alpha, int alpha_per_image,
                            const int32_t *cand,
                            int K, int32_t *best_idx, double *best_d64, float *d32, double *d64,
                            const int32_t *image_order, void *stream)
{
    MTIP_REQUIRE(K == 1 || K == 2 || K == 4 || K == 8, MTIP_ERR_ARG, "mtip_rescore: K must be 1, 2, 4 or 8");
    MTIP_REQUIRE(N_i >= 0 && N_p >= 0, MTIP_ERR_ARG, "mtip_rescore: negative size");
    if (N_i == 0) return MTIP_OK;
    MTIP_REQUIRE(N_i <= 0x7fffffff, MTIP_ERR_ARG, "mtip_rescore: N_i too large");
    MTIP_REQUIRE(slices && images && alpha && cand, MTIP_ERR_ARG, "mtip_rescore: null pointer");
    MTIP_REQUIRE(ld_s >= N_p && ld_d >= N_p, MTIP_ERR_ARG, "mtip_rescore: leading dimension < N_p");
    hipStream_t st = (hipStream_t)stream;
    dim3 g((unsigned)N_i);
    switch (K) {
    case 1: hipLaunchKernelGGL(k_rescore<1>, g, dim3(256), 0, st, slices, ld_s, N_o, images, ld_d, N_p, alpha, alpha_per_image, cand, best_idx, best_d64, d32, d64, image_order); break;
    case 2: hipLaunchKernelGGL(k_rescore<2>, g, dim3(256), 0, st, slices, ld_s, N_o, images, ld_d, N_p, alpha, alpha_per_image, cand, best_idx, best_d64, d32, d64, image_order); break;
    case 4: hipLaunchKernelGGL(k_rescore<4>, g, dim3(256), 0, st, slices, ld_s, N_o, images, ld_d, N_p, alpha, alpha_per_image, cand, best_idx, best_d64, d32, d64, image_order); break;
    default: hipLaunchKernelGGL(k_rescore<8>, g, dim3(256), 0, st, slices, ld_s, N_o, images, ld_d, N_p, alpha, alpha_per_image, cand, best_idx, best_d64, d32, d64, image_order); break;
    }
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}

extern "C" int mtip_row_stats(const float *in, int64_t rows, int64_t cols, int64_t ld,
                              double *row_sum, double *row_sumsq, void *stream)
{
    MTIP_REQUIRE(rows >= 0 && cols >= 0 && ld >= cols, MTIP_ERR_ARG, "mtip_row_stats: bad shape");
    if (rows == 0) return MTIP_OK;
    MTIP_REQUIRE(rows <= 0x7fffffff, MTIP_ERR_ARG, "mtip_row_stats: too many rows");
    MTIP_REQUIRE(in && (row_sum || row_sumsq), MTIP_ERR_ARG, "mtip_row_stats: null pointer");
    hipLaunchKernelGGL(k_row_stats, dim3((unsigned)rows), dim3(256), 0, (hipStream_t)stream, in,
                       cols, ld, row_sum, row_sumsq);
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}

extern "C" int mtip_scaling_ratio(const double *img_sum, const double *img_sumsq, int64_t n_img,
                                  const double *sl_sum, const double *sl_sumsq, int64_t n_sl,
                                  int64_t n_pix, double *alpha, void *stream)
{
    MTIP_REQUIRE(n_img > 0 && n_sl > 0 && n_pix > 0, MTIP_ERR_ARG, "mtip_scaling_ratio: empty input");
    MTIP_REQUIRE(img_sum && img_sumsq && sl_sum && sl_sumsq && alpha, MTIP_ERR_ARG,
                 "mtip_scaling_ratio: null pointer");
    hipLaunchKernelGGL(k_scaling_ratio, dim3(1), dim3(256), 0, (hipStream_t)stream, img_sum,
                       img_sumsq, n_img, sl_sum, sl_sumsq, n_sl, n_pix, alpha);
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}

extern "C" int mtip_residual_rows(const float *in, int64_t rows, int64_t cols, int64_t ld_in,
                                  const double *mult, int mult_flags, const float *center,
                                  int format, void *out, int64_t ld_out, float *scales,
                                  double *sumsq, void *stream)
{
    MTIP_REQUIRE(rows >= 0 && cols >= 0 && ld_in >= cols && ld_out >= cols, MTIP_ERR_ARG,
                 "mtip_residual_rows: bad shape");
    if (rows == 0) return MTIP_OK;
    MTIP_REQUIRE(rows <= 0x7fffffff, MTIP_ERR_ARG, "mtip_residual_rows: too many rows");
    MTIP_REQUIRE(in && out, MTIP_ERR_ARG, "mtip_residual_rows: null pointer");
    hipStream_t st = (hipStream_t)stream;
    dim3 g((unsigned)rows);
    // 16-byte-aligned rows (the engine's pitches and allocations always are)
    const int vec = ((uintptr_t)in % 16 == 0) && ld_in % 4 == 0 && ld_out % 4 == 0 &&
                    ((uintptr_t)out % 16 == 0) && (!center || (uintptr_t)center % 16 == 0);
    if (format == MTIP_FP32)
        hipLaunchKernelGGL(k_residual<MTIP_FP32>, g, dim3(256), 0, st, in, cols, ld_in, mult,
                           mult_flags, center, out, ld_out, scales, sumsq, vec);
    else if (format == MTIP_BF16)
        hipLaunchKernelGGL(k_residual<MTIP_BF16>, g, dim3(256), 0, st, in, cols, ld_in, mult,
                           mult_flags, center, out, ld_out, scales, sumsq, vec);
    else if (format == MTIP_FP8) {
        MTIP_REQUIRE(scales, MTIP_ERR_ARG, "mtip_residual_rows: FP8 needs a scales array");
        hipLaunchKernelGGL(k_residual<MTIP_FP8>, g, dim3(256), 0, st, in, cols, ld_in, mult,
                           mult_flags, center, out, ld_out, scales, sumsq, vec);
    } else {
        return mtip::fail(MTIP_ERR_ARG, "mtip_residual_rows: unknown format");
    }
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}

extern "C" int mtip_quantize_rows(const float *in, int64_t rows, int64_t cols, int64_t ld_in,
                                  int format, void *out, int64_t ld_out, float *scales,
                                  void *stream)
{
    MTIP_REQUIRE(format == MTIP_BF16 || format == MTIP_FP8, MTIP_ERR_ARG,
                 "mtip_quantize_rows: format must be MTIP_BF16 or MTIP_FP8");
    return mtip_residual_rows(in, rows, cols, ld_in, nullptr, 0, nullptr, format, out, ld_out,
                              scales, nullptr, stream);
}

extern "C" size_t mtip_col_mean_workspace_size(int64_t cols)
{
    return cols > 0 ? (size_t)cols * kColGroups * sizeof(double) : 0;
}

extern "C" int mtip_col_mean(const float *in, int64_t rows, int64_t cols, int64_t ld, double *work,
                             size_t work_bytes, float *mean, void *stream)
{
    MTIP_REQUIRE(rows > 0 && cols >= 0 && ld >= cols, MTIP_ERR_ARG, "mtip_col_mean: bad shape");
    if (cols == 0) return MTIP_OK;
    MTIP_REQUIRE(in && mean && work, MTIP_ERR_ARG, "mtip_col_mean: null pointer");
    MTIP_REQUIRE(work_bytes >= mtip_col_mean_workspace_size(cols), MTIP_ERR_NOMEM,
                 "mtip_col_mean: workspace too small");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_col_partial, dim3((unsigned)mtip::ceil_div(cols, 64), kColGroups),
                       dim3(256), 0, st, in, rows, cols, ld, work);
    hipLaunchKernelGGL(k_col_finish, dim3((unsigned)mtip::ceil_div(cols, 256)), dim3(256), 0, st,
                       work, rows, cols, mean);
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}

extern "C" int mtip_euclidean_distance(const float *ref, const float *data, int64_t numRef,
                                       int64_t numData, int64_t totalPixels, float *euDist,
                                       void *stream)
{
    MTIP_REQUIRE(numRef >= 0 && numData >= 0 && totalPixels >= 0, MTIP_ERR_ARG,
                 "mtip_euclidean_distance: negative size");
    if (numRef == 0 || numData == 0) return MTIP_OK;
    // rows are staged in 16-byte chunks that lie wholly inside or wholly
    // past a row (the part past it reads as zero), so a row must be a whole
    // number of chunks
    MTIP_REQUIRE(totalPixels > 0 && totalPixels % 4 == 0, MTIP_ERR_ARG,
                 "mtip_euclidean_distance: totalPixels must be a positive multiple of 4");
    MTIP_REQUIRE(ref && data && euDist, MTIP_ERR_ARG, "mtip_euclidean_distance: null pointer");
    MTIP_REQUIRE(((uintptr_t)ref % 16) == 0 && ((uintptr_t)data % 16) == 0, MTIP_ERR_ARG,
                 "mtip_euclidean_distance: inputs must be 16-byte aligned");
    const int nta = (int)mtip::ceil_div(numRef, BM);
    const int64_t ntb = mtip::ceil_div(numData, BN);
    MTIP_REQUIRE((int64_t)nta * ntb < 0x7fffffff, MTIP_ERR_ARG, "mtip_euclidean_distance: grid too large");
    hipLaunchKernelGGL((k_gemm<MTIP_FP32, 1, OUT_FULL>), dim3((unsigned)(nta * ntb)), dim3(256), 0,
                       (hipStream_t)stream, (const uint8_t *)ref, totalPixels * 4, numRef,
                       (const uint8_t *)data, totalPixels * 4, numData, totalPixels * 4, nullptr,
                       nullptr, nullptr, nullptr, nta, nullptr, euDist, numRef);
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}

extern "C" int mtip_heap_sort(const float *euDist, int64_t numData, int64_t numRef, int totalNN,
                              int32_t *idx, void *stream)
{
    MTIP_REQUIRE(numData >= 0 && numRef >= 0, MTIP_ERR_ARG, "mtip_heap_sort: negative size");
    MTIP_REQUIRE(totalNN >= 1, MTIP_ERR_ARG, "mtip_heap_sort: totalNN must be at least 1");
    if (numData == 0) return MTIP_OK;
    MTIP_REQUIRE(numRef <= 0x7fffffff, MTIP_ERR_ARG, "mtip_heap_sort: numRef too large");
    MTIP_REQUIRE(numData <= 0x7fffffff, MTIP_ERR_ARG, "mtip_heap_sort: numData too large");
    MTIP_REQUIRE(euDist && idx, MTIP_ERR_ARG, "mtip_heap_sort: null pointer");
    hipStream_t st = (hipStream_t)stream;
    if (totalNN > 8) {
        hipLaunchKernelGGL(k_select_rows, dim3((unsigned)numData), dim3(256), 0, st, euDist, numData,
                           numRef, numRef, (int64_t)totalNN, idx);
    } else if (totalNN == 1)
        launch_topk_rows<1>(nullptr, euDist, numData, numRef, numRef, totalNN, idx, nullptr, st);
    else if (totalNN == 2)
        launch_topk_rows<2>(nullptr, euDist, numData, numRef, numRef, totalNN, idx, nullptr, st);
    else if (totalNN <= 4)
        launch_topk_rows<4>(nullptr, euDist, numData, numRef, numRef, totalNN, idx, nullptr, st);
    else
        launch_topk_rows<8>(nullptr, euDist, numData, numRef, numRef, totalNN, idx, nullptr, st);
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}

extern "C" int mtip_segment_scaling_ratio(const double *img_sum, const double *img_sumsq,
                                          const int64_t *seg_offsets, int64_t n_seg,
                                          const double *sl_sum, const double *sl_sumsq, int64_t n_sl,
                                          int64_t n_pix, double *alpha, void *stream)
{
    MTIP_REQUIRE(n_seg >= 0 && n_sl > 0 && n_pix > 0, MTIP_ERR_ARG, "mtip_segment_scaling_ratio: bad size");
    if (n_seg == 0) return MTIP_OK;
    MTIP_REQUIRE(n_seg <= 0x7fffffff, MTIP_ERR_ARG, "mtip_segment_scaling_ratio: too many segments");
    MTIP_REQUIRE(img_sum && img_sumsq && seg_offsets && sl_sum && sl_sumsq && alpha, MTIP_ERR_ARG,
                 "mtip_segment_scaling_ratio: null pointer");
    hipLaunchKernelGGL(k_segment_ratio, dim3((unsigned)n_seg), dim3(256), 0, (hipStream_t)stream,
                       img_sum, img_sumsq, seg_offsets, sl_sum, sl_sumsq, n_sl, n_pix, alpha);
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}
