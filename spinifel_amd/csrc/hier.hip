// Batched matching against a hierarchy of orientation grids (gfx950).
//
// After a dense top-K against the coarse grid, every image refines around
// its winner on the next grid: the candidates are the winner's children (the
// finer nodes that have the winner among their nearest coarser nodes).
// Images are grouped by winner on the device (count, scan, scatter, task
// list), so the images of one group share one read of the children's slices;
// a persistent kernel then walks the (group, 64-entry chunk) tasks.  An
// image may follow several winners (a beam of T): grouping then works on
// entries (image, t), each entry gets its own top-K, and a merge keeps each
// image's best distinct candidates.  No
// step needs the host, so a whole multi-level pass can be captured in a
// hipGraph and replayed.
//
// Grouped top-K block: per wave 32 children (MFMA A rows) x 2 x 32 entries
// (B rows) on v_mfma_f32_32x32x16_bf16.  Each lane streams its own operand
// rows with 16-byte loads (rows are gathered, so there is nothing to share
// through LDS).  Within each 64-element chunk of pixels, lane half h feeds
// elements [32 h, 32 h + 32) in four MFMA steps, for both operands alike,
// so every pixel meets itself.
#include "mtip_internal.h"

namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int intx8 __attribute__((ext_vector_type(8)));

constexpr int kTaskImages = 64;     // entries per task (two MFMA blocks of columns)
constexpr uint64_t KEY_NONE = ~0ull;

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Group sizes over entries e = i * n_win + t (image i, its t-th winner):
// counts[g] += 1 for every entry whose winner is g.
__global__ void k_count(const int32_t *__restrict__ win, int64_t win_stride, int64_t n_win,
                        int64_t N_e, int64_t n_groups, int32_t *__restrict__ counts,
                        int *__restrict__ bad)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t g = e < N_e ? win[(e / n_win) * win_stride + e % n_win] : 0;
    const bool ok = g >= 0 && g < n_groups;
    if (e < N_e && ok) atomicAdd(&counts[g], 1);
    // entries left out (often many: repeats marked -1) add to one counter,
    // once per wave rather than once per entry
    const uint64_t m = __ballot(e < N_e && !ok);
    if (m && (threadIdx.x & 63) == (int)__ffsll((unsigned long long)m) - 1)
        atomicAdd(bad, (int)__popcll(m));
}

// One workgroup: exclusive scans of the group sizes (image offsets) and of
// the per-group task counts; cursor = image offsets (for the scatter);
// n_tasks[0] = number of tasks.
__global__ __launch_bounds__(1024) void k_scan(const int32_t *__restrict__ counts, int64_t n_groups,
                                               int32_t *__restrict__ img_off,
                                               int32_t *__restrict__ cursor,
                                               int32_t *__restrict__ task_off,
                                               int32_t *__restrict__ n_tasks)
{
    __shared__ int32_t s_a[1024], s_b[1024];
    const int tid = threadIdx.x;
    const int64_t per = (n_groups + 1023) / 1024;
    const int64_t g0 = tid * per, g1 = min(n_groups, g0 + per);
    int32_t a = 0, b = 0;
    for (int64_t g = g0; g < g1; ++g) {
        a += counts[g];
        b += (counts[g] + kTaskImages - 1) / kTaskImages;
    }
    s_a[tid] = a;
    s_b[tid] = b;
    __syncthreads();
    // Hillis-Steele inclusive scan over the 1024 partial sums
    for (int d = 1; d < 1024; d <<= 1) {
        int32_t va = tid >= d ? s_a[tid - d] : 0, vb = tid >= d ? s_b[tid - d] : 0;
        __syncthreads();
        s_a[tid] += va;
        s_b[tid] += vb;
        __syncthreads();
    }
    int32_t ra = s_a[tid] - a, rb = s_b[tid] - b;     // exclusive
    for (int64_t g = g0; g < g1; ++g) {
        img_off[g] = ra;
        cursor[g] = ra;
        task_off[g] = rb;
        ra += counts[g];
        rb += (counts[g] + kTaskImages - 1) / kTaskImages;
    }
    if (tid == 1023) {
        img_off[n_groups] = s_a[1023];
        cursor[n_groups] = s_a[1023];     // entries left out go after the grouped ones
        n_tasks[0] = s_b[1023];
    }
}

// perm[pos] = entry, the entries of a group contiguous (order within a group
// is arbitrary; every per-entry result is independent of it), entries with
// no valid winner at the end: perm is a permutation of all entries.
__global__ void k_scatter(const int32_t *__restrict__ win, int64_t win_stride, int64_t n_win,
                          int64_t N_e, int64_t n_groups, int32_t *__restrict__ cursor,
                          int32_t *__restrict__ perm)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t g = e < N_e ? win[(e / n_win) * win_stride + e % n_win] : 0;
    const bool ok = g >= 0 && g < n_groups;
    if (e < N_e && ok) perm[atomicAdd(&cursor[g], 1)] = (int32_t)e;
    // the entries left out share one cursor: one atomic per wave, each lane
    // placed by its rank among the wave's left-out entries
    const uint64_t m = __ballot(e < N_e && !ok);
    if (m) {
        const int lane = threadIdx.x & 63, leader = (int)__ffsll((unsigned long long)m) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(&cursor[n_groups], (int)__popcll(m));
        base = __shfl(base, leader, 64);
        const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (e < N_e && !ok) perm[base + rank] = (int32_t)e;
    }
}

// tasks[task_off[g] + t] = (g, first image position, end position)
__global__ void k_tasks(const int32_t *__restrict__ counts, const int32_t *__restrict__ img_off,
                        const int32_t *__restrict__ task_off, int64_t n_groups,
                        int4 *__restrict__ tasks)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_groups) return;
    const int32_t n = counts[g], b = img_off[g];
    int32_t t = task_off[g];
    for (int32_t s = 0; s < n; s += kTaskImages, ++t)
        tasks[t] = make_int4((int32_t)g, b + s, b + min(n, s + kTaskImages), 0);
}

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v)
{
    return __builtin_bit_cast(bf16x8, v);
}

__device__ __forceinline__ intx8 as_intx8(uint4 lo, uint4 hi)
{
    return intx8{(int)lo.x, (int)lo.y, (int)lo.z, (int)lo.w, (int)hi.x, (int)hi.y, (int)hi.z, (int)hi.w};
}

// Per task: the K children of the group with the smallest
//   v = ||s_c||^2 - 2 s_c . d_e
// for each of up to 64 entries (the centred, scaled operands of the dense
// path, so v ranks like the squared distance).  BF16 operands, or e4m3 with
// per-row scales (then s_c . d_e = scale_c scale_e (q_c . q_e)).  kd_b:
// bytes summed per row, a multiple of 64 within both row pitches; columns
// past N_p are zero.
//
// Wave w takes children c0 + 32 w .. c0 + 32 w + 31 of each pass of 128,
// for two 32-entry column blocks, reading its children's rows straight from
// memory (each child row once per task).  The 64 entries' rows are shared
// by the four waves: each chunk of kChunkBytes of them is staged once in
// LDS (double-buffered), so an entry row is read from memory once per pass
// instead of once per wave.  A chunk's loads (children and entries) are
// issued a whole chunk ahead of their MFMAs: the kernel is bound by that
// memory latency, so a chunk is 256 bytes of each row (two BF16 or four e4m3
// MFMA steps).  The waves' running top-K lists are merged in LDS at the end.
constexpr int kXcdsH = 8;                      // XCDs of the device (workgroups dealt round robin)
constexpr int kChunkBytes = 256;               // bytes of each row per chunk
constexpr int kEntStrideB = kChunkBytes + 16;  // bytes per staged entry row
constexpr int kEntPieces = kTaskImages * kChunkBytes / 16 / 256;   // 16-byte pieces per thread

__device__ __forceinline__ uint4 ld16b(const uint8_t *row, int64_t e)
{
    return row ? *(const uint4 *)(row + e) : make_uint4(0, 0, 0, 0);
}

template <int PREC, int K>
__global__ __launch_bounds__(256) void k_group_topk(
    const int4 *__restrict__ tasks, const int32_t *__restrict__ n_tasks,
    const int32_t *__restrict__ perm, const int32_t *__restrict__ child_off,
    const int32_t *__restrict__ child_idx, const uint8_t *__restrict__ S, int64_t lds_b,
    const float *__restrict__ s_scale, const double *__restrict__ s_sumsq,
    const uint8_t *__restrict__ D, int64_t ldd_b, const float *__restrict__ d_scale,
    int64_t kd_b, int n_win, int32_t *__restrict__ out_idx, float *__restrict__ out_val)
{
    __shared__ uint64_t s_top[4][kTaskImages][K];
    __shared__ __attribute__((aligned(16))) uint8_t s_ent[2][kTaskImages * kEntStrideB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int half = lane >> 5, l32 = lane & 31;
    const int nt = n_tasks[0];
    // staging: piece q of thread tid is entry row (tid + 256 q) / 16, bytes
    // 16 ((tid + 256 q) % 16) .. + 15 of a chunk
    constexpr int kRowPieces = kChunkBytes / 16;
    int st_row[kEntPieces], st_col[kEntPieces];
#pragma unroll
    for (int q = 0; q < kEntPieces; ++q) {
        st_row[q] = (tid + 256 * q) / kRowPieces;
        st_col[q] = ((tid + 256 * q) % kRowPieces) * 16;
    }
    // the byte offset of this lane's s-th 16-byte piece of a child row in a
    // chunk: BF16, two 128-byte MFMA steps, each lane half 64 bytes in four
    // pieces; e4m3, four 64-byte steps, each lane half 32 bytes in two
    auto a_off = [&](int s) __attribute__((always_inline)) {
        return PREC == MTIP_BF16 ? 128 * (s >> 2) + 64 * half + 16 * (s & 3)
                                 : 64 * (s >> 1) + 32 * half + 16 * (s & 1);
    };
    // kd_b is a multiple of 64; a chunk past it reads zeros
    const int64_t nk = (kd_b + kChunkBytes - 1) / kChunkBytes;

    // workgroup b runs on XCD b % 8: XCD x takes the x-th of 8 contiguous
    // task ranges, dealt round robin to its workgroups, so the tasks of one
    // group (adjacent in the list) run together on one XCD and read their
    // children's rows through one L2
    const int xcd = blockIdx.x % kXcdsH, local = blockIdx.x / kXcdsH;
    const int xcd_blocks = ((int)gridDim.x - xcd + kXcdsH - 1) / kXcdsH;
    const int t_end = (int)((int64_t)nt * (xcd + 1) / kXcdsH);
    for (int task = (int)((int64_t)nt * xcd / kXcdsH) + local; task < t_end;
         task += xcd_blocks) {
        const int4 tk = tasks[task];
        const int g = tk.x, p0 = tk.y, p1 = tk.z;
        const int c_begin = child_off[g], c_end = child_off[g + 1];
        const uint8_t *erow[kEntPieces];
#pragma unroll
        for (int q = 0; q < kEntPieces; ++q) {
            const int pe = p0 + st_row[q];
            erow[q] = pe < p1 ? D + (int64_t)(perm[pe] / n_win) * ldd_b : nullptr;
        }
        // e4m3: the scales of this lane's two entry columns
        float esc[2] = {1.0f, 1.0f};
        if constexpr (PREC == MTIP_FP8) {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int pe = p0 + 32 * j + l32;
                esc[j] = pe < p1 ? d_scale[perm[pe] / n_win] : 0.0f;
            }
        }
        uint64_t top[2][K];
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < K; ++r) top[j][r] = KEY_NONE;

        // bytes e.. of a row, or zero past kd_b (or for a missing row)
        auto ldk = [&](const uint8_t *row, int64_t e) __attribute__((always_inline)) {
            return e < kd_b ? ld16b(row, e) : make_uint4(0, 0, 0, 0);
        };
        for (int cp = c_begin; cp < c_end; cp += 128) {
            const int c0 = cp + 32 * wave;
            const int crow = c0 + l32;
            const int child = crow < c_end ? child_idx[crow] : -1;
            const uint8_t *srow = child >= 0 ? S + (int64_t)child * lds_b : nullptr;
            floatx16 acc[2];
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int v = 0; v < 16; ++v) acc[j][v] = 0.0f;
            // chunk 0: entries into LDS buffer 0, this wave's children into registers
            uint4 a[8], en[kEntPieces];
#pragma unroll
            for (int q = 0; q < kEntPieces; ++q) en[q] = ldk(erow[q], st_col[q]);
#pragma unroll
            for (int q = 0; q < kEntPieces; ++q)
                *(uint4 *)(s_ent[0] + st_row[q] * kEntStrideB + st_col[q]) = en[q];
#pragma unroll
            for (int s = 0; s < 8; ++s) a[s] = ldk(srow, a_off(s));
            __syncthreads();
            for (int64_t kc = 0; kc < nk; ++kc) {
                const bool more = kc + 1 < nk;
                const int64_t e1 = (kc + 1) * kChunkBytes;
                uint4 an[8];
#pragma unroll
                for (int q = 0; q < kEntPieces; ++q)
                    en[q] = more ? ldk(erow[q], e1 + st_col[q]) : make_uint4(0, 0, 0, 0);
#pragma unroll
                for (int s = 0; s < 8; ++s)
                    an[s] = more ? ldk(srow, e1 + a_off(s)) : make_uint4(0, 0, 0, 0);
                const uint8_t *eb = s_ent[kc & 1];
                if (c0 < c_end) {       // wave-uniform: a wave past the children only stages
                    if constexpr (PREC == MTIP_BF16) {
#pragma unroll
                        for (int s = 0; s < 8; ++s) {
                            const int off = a_off(s);
                            const uint4 b0 = *(const uint4 *)(eb + l32 * kEntStrideB + off);
                            const uint4 b1 = *(const uint4 *)(eb + (32 + l32) * kEntStrideB + off);
                            acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                                as_bf16x8(a[s]), as_bf16x8(b0), acc[0], 0, 0, 0);
                            acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                                as_bf16x8(a[s]), as_bf16x8(b1), acc[1], 0, 0, 0);
                        }
                    } else {
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
                            const int off = a_off(2 * t);
                            const intx8 av = as_intx8(a[2 * t], a[2 * t + 1]);
                            const uint8_t *r0 = eb + l32 * kEntStrideB + off;
                            const uint8_t *r1 = eb + (32 + l32) * kEntStrideB + off;
                            const intx8 b0 = as_intx8(*(const uint4 *)r0, *(const uint4 *)(r0 + 16));
                            const intx8 b1 = as_intx8(*(const uint4 *)r1, *(const uint4 *)(r1 + 16));
                            acc[0] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                                av, b0, acc[0], 0, 0, 0, 127, 0, 127);
                            acc[1] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                                av, b1, acc[1], 0, 0, 0, 127, 0, 127);
                        }
                    }
                }
                // the other buffer was last read in chunk kc - 1, before the
                // barrier that ended it
                if (more) {
                    uint8_t *nb = s_ent[(kc + 1) & 1];
#pragma unroll
                    for (int q = 0; q < kEntPieces; ++q)
                        *(uint4 *)(nb + st_row[q] * kEntStrideB + st_col[q]) = en[q];
                }
#pragma unroll
                for (int s = 0; s < 8; ++s) a[s] = an[s];
                __syncthreads();
            }
            // rows of this lane: child c0 + 8 (v / 4) + 4 half + v % 4
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int r = c0 + 8 * (v >> 2) + 4 * half + (v & 3);
                if (r < c_end) {
                    const int cid = child_idx[r];
                    const float ss = (float)s_sumsq[cid];
                    const float csc = PREC == MTIP_FP8 ? s_scale[cid] : 1.0f;
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        const float dot = PREC == MTIP_FP8 ? acc[j][v] * (csc * esc[j]) : acc[j][v];
                        mtip::topk_insert<K>(top[j], mtip::make_key(fmaf(-2.0f, dot, ss),
                                                                    (uint32_t)cid));
                    }
                }
            }
        }
        // merge the lane halves, then the four waves through LDS
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            uint64_t other[K];
#pragma unroll
            for (int r = 0; r < K; ++r) other[r] = mtip::shfl_xor_u64(top[j][r], 32);
            mtip::topk_merge<K>(top[j], other);
            if (half == 0)
#pragma unroll
                for (int r = 0; r < K; ++r) s_top[wave][32 * j + l32][r] = top[j][r];
        }
        __syncthreads();
        if (tid < kTaskImages && p0 + tid < p1) {
            uint64_t t[K], o[K];
#pragma unroll
            for (int r = 0; r < K; ++r) t[r] = s_top[0][tid][r];
#pragma unroll
            for (int w = 1; w < 4; ++w) {
#pragma unroll
                for (int r = 0; r < K; ++r) o[r] = s_top[w][tid][r];
                mtip::topk_merge<K>(t, o);
            }
            const int64_t ent = perm[p0 + tid];     // the entry (its own output row)
#pragma unroll
            for (int r = 0; r < K; ++r) {
                const bool ok = t[r] != KEY_NONE;
                out_idx[ent * K + r] = ok ? (int32_t)(uint32_t)t[r] : -1;
                out_val[ent * K + r] =
                    ok ? mtip::ordered_to_float((uint32_t)(t[r] >> 32)) : __builtin_inff();
            }
        }
        __syncthreads();
    }
}

// Per image: the K_out smallest distinct keys among its n_in candidate
// (index, value) pairs (T entries x K each, index -1 = none).
template <int KO>
__global__ void k_merge_distinct(const int32_t *__restrict__ in_idx,
                                 const float *__restrict__ in_val, int64_t N_i, int n_in,
                                 int32_t *__restrict__ out_idx, float *__restrict__ out_val)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N_i) return;
    uint64_t top[KO];
#pragma unroll
    for (int r = 0; r < KO; ++r) top[r] = KEY_NONE;
    for (int j = 0; j < n_in; ++j) {
        const int32_t id = in_idx[i * n_in + j];
        if (id < 0) continue;
        const uint64_t key = mtip::make_key(in_val[i * n_in + j], (uint32_t)id);
        bool dup = false;
#pragma unroll
        for (int r = 0; r < KO; ++r) dup |= (top[r] == key);
        if (!dup) mtip::topk_insert<KO>(top, key);
    }
#pragma unroll
    for (int r = 0; r < KO; ++r) {
        const bool ok = top[r] != KEY_NONE;
        out_idx[i * KO + r] = ok ? (int32_t)(uint32_t)top[r] : -1;
        out_val[i * KO + r] = ok ? mtip::ordered_to_float((uint32_t)(top[r] >> 32)) : __builtin_inff();
    }
}

// Workgroups of k_group_topk<K> that fit on the device at once (the
// persistent grid: every one of them is resident from the start)
template <int PREC>
int n_blocks_persistent(int device, int K)
{
    // one query per device, precision and K (the answer does not change; a
    // race writes the same value)
    static int cached[16][4];
    const int ks = K == 1 ? 0 : K == 2 ? 1 : K == 4 ? 2 : 3;
    if (device >= 0 && device < 16 && cached[device][ks] > 0) return cached[device][ks];
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        cus <= 0)
        cus = 256;
    int per_cu = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (K) {
    case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_group_topk<PREC, 1>, 256, 0); break;
    case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_group_topk<PREC, 2>, 256, 0); break;
    case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_group_topk<PREC, 4>, 256, 0); break;
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_group_topk<PREC, 8>, 256, 0); break;
    }
    if (e != hipSuccess || per_cu <= 0) per_cu = 2;
    if (device >= 0 && device < 16) cached[device][ks] = per_cu * cus;
    return per_cu * cus;
}

}  // namespace

extern "C" size_t mtip_group_workspace_size(int64_t N_e, int64_t n_groups)
{
    // counts, img_off (+1), cursor, task_off, n_tasks, bad | perm | tasks (int4)
    const int64_t ints = 4 * (n_groups + 1) + 2 + N_e;
    const int64_t max_tasks = N_e / kTaskImages + n_groups + 1;
    return (size_t)(((ints * 4 + 15) / 16) * 16 + max_tasks * 16);
}

namespace {
struct GroupWs {
    int32_t *counts, *img_off, *cursor, *task_off, *n_tasks, *bad, *perm;
    int4 *tasks;
};
GroupWs carve(void *ws, int64_t N_i, int64_t n_groups)
{
    GroupWs w;
    int32_t *p = (int32_t *)ws;
    w.counts = p;
    p += n_groups + 1;
    w.img_off = p;
    p += n_groups + 1;
    w.cursor = p;
    p += n_groups + 1;
    w.task_off = p;
    p += n_groups + 1;
    w.n_tasks = p++;
    w.bad = p++;
    w.perm = p;
    p += N_i;
    const int64_t ints = 4 * (n_groups + 1) + 2 + N_i;
    w.tasks = (int4 *)((uint8_t *)ws + ((ints * 4 + 15) / 16) * 16);
    return w;
}
}  // namespace

extern "C" size_t mtip_group_perm_offset(int64_t N_e, int64_t n_groups)
{
    (void)N_e;
    return (size_t)(4 * (n_groups + 1) + 2) * sizeof(int32_t);
}

extern "C" int mtip_group_by_winner(const int32_t *winners, int64_t win_stride, int n_win,
                                    int64_t N_i, int64_t n_groups, void *workspace,
                                    size_t ws_bytes, void *stream)
{
    MTIP_REQUIRE(N_i >= 0 && n_groups > 0 && n_win >= 1 && win_stride >= n_win, MTIP_ERR_ARG,
                 "mtip_group_by_winner: bad sizes");
    const int64_t N_e = N_i * n_win;
    MTIP_REQUIRE(N_e <= 0x7fffffff && n_groups <= 0x7fffffff, MTIP_ERR_ARG,
                 "mtip_group_by_winner: sizes exceed int32");
    MTIP_REQUIRE(workspace, MTIP_ERR_ARG, "mtip_group_by_winner: null workspace");
    MTIP_REQUIRE(ws_bytes >= mtip_group_workspace_size(N_e, n_groups), MTIP_ERR_ARG,
                 "mtip_group_by_winner: workspace too small");
    MTIP_REQUIRE(N_i == 0 || winners, MTIP_ERR_ARG, "mtip_group_by_winner: null winners");
    hipStream_t st = (hipStream_t)stream;
    GroupWs w = carve(workspace, N_e, n_groups);
    // counts, n_tasks and bad start at zero (a kernel, not a memset, so a
    // captured graph can replay it)
    const int64_t nz = 4 * (n_groups + 1) + 2;
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)mtip::ceil_div(nz, 256)), dim3(256), 0, st,
                       w.counts, nz, 0);
    if (N_e > 0)
        hipLaunchKernelGGL(k_count, dim3((unsigned)mtip::ceil_div(N_e, 256)), dim3(256), 0, st,
                           winners, win_stride, (int64_t)n_win, N_e, n_groups, w.counts, w.bad);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, w.counts, n_groups, w.img_off,
                       w.cursor, w.task_off, w.n_tasks);
    if (N_e > 0)
        hipLaunchKernelGGL(k_scatter, dim3((unsigned)mtip::ceil_div(N_e, 256)), dim3(256), 0, st,
                           winners, win_stride, (int64_t)n_win, N_e, n_groups, w.cursor, w.perm);
    hipLaunchKernelGGL(k_tasks, dim3((unsigned)mtip::ceil_div(n_groups, 256)), dim3(256), 0, st,
                       w.counts, w.img_off, w.task_off, n_groups, w.tasks);
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}

namespace {

int group_topk(const char *fn, int precision, const void *workspace, int n_win, int64_t N_i,
               int64_t n_groups, const int32_t *child_off, const int32_t *child_idx,
               const void *slices, int64_t ld_s, const float *slice_scale,
               const double *slice_sumsq, const void *images, int64_t ld_d,
               const float *image_scale, int64_t kd, int K, int32_t *top_idx, float *top_val,
               int device, void *stream)
{
    MTIP_REQUIRE(precision == MTIP_BF16 || precision == MTIP_FP8, MTIP_ERR_ARG,
                 std::string(fn) + ": precision must be BF16 or FP8");
    MTIP_REQUIRE(N_i >= 0 && n_groups > 0 && n_win >= 1, MTIP_ERR_ARG,
                 std::string(fn) + ": bad sizes");
    MTIP_REQUIRE(K == 1 || K == 2 || K == 4 || K == 8, MTIP_ERR_ARG,
                 std::string(fn) + ": K must be 1, 2, 4 or 8");
    MTIP_REQUIRE(kd > 0 && kd % 64 == 0, MTIP_ERR_ARG,
                 std::string(fn) + ": kd must be a positive multiple of 64");
    const int eb = precision == MTIP_FP8 ? 1 : 2;      // bytes per element
    MTIP_REQUIRE(ld_s >= kd && ld_d >= kd && (ld_s * eb) % 16 == 0 && (ld_d * eb) % 16 == 0,
                 MTIP_ERR_ARG,
                 std::string(fn) + ": row pitches must be >= kd and whole 16-byte units");
    if (N_i == 0) return MTIP_OK;
    MTIP_REQUIRE(workspace && child_off && child_idx && slices && slice_sumsq && images &&
                     top_idx && top_val,
                 MTIP_ERR_ARG, std::string(fn) + ": null pointer");
    MTIP_REQUIRE(precision != MTIP_FP8 || (slice_scale && image_scale), MTIP_ERR_ARG,
                 std::string(fn) + ": e4m3 operands need their row scales");
    hipStream_t st = (hipStream_t)stream;
    GroupWs w = carve(const_cast<void *>(workspace), N_i * n_win, n_groups);
#define MTIP_GROUP_CASE(PP, KK)                                                                \
    hipLaunchKernelGGL((k_group_topk<PP, KK>), dim3(n_blocks_persistent<PP>(device, KK)),    \
                       dim3(256), 0, st, w.tasks, w.n_tasks, w.perm, child_off, child_idx,   \
                       (const uint8_t *)slices, ld_s * eb, slice_scale, slice_sumsq,          \
                       (const uint8_t *)images, ld_d * eb, image_scale, kd * eb, n_win,       \
                       top_idx, top_val)
    if (precision == MTIP_BF16) {
        switch (K) {
        case 1: MTIP_GROUP_CASE(MTIP_BF16, 1); break;
        case 2: MTIP_GROUP_CASE(MTIP_BF16, 2); break;
        case 4: MTIP_GROUP_CASE(MTIP_BF16, 4); break;
        default: MTIP_GROUP_CASE(MTIP_BF16, 8); break;
        }
    } else {
        switch (K) {
        case 1: MTIP_GROUP_CASE(MTIP_FP8, 1); break;
        case 2: MTIP_GROUP_CASE(MTIP_FP8, 2); break;
        case 4: MTIP_GROUP_CASE(MTIP_FP8, 4); break;
        default: MTIP_GROUP_CASE(MTIP_FP8, 8); break;
        }
    }
#undef MTIP_GROUP_CASE
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}

}  // namespace

extern "C" int mtip_group_topk(const void *workspace, int n_win, int64_t N_i, int64_t n_groups,
                               const int32_t *child_off, const int32_t *child_idx,
                               const uint16_t *slices, int64_t ld_s, const double *slice_sumsq,
                               const uint16_t *images, int64_t ld_d, int64_t kd, int K,
                               int32_t *top_idx, float *top_val, int device, void *stream)
{
    return group_topk("mtip_group_topk", MTIP_BF16, workspace, n_win, N_i, n_groups, child_off,
                      child_idx, slices, ld_s, nullptr, slice_sumsq, images, ld_d, nullptr, kd, K,
                      top_idx, top_val, device, stream);
}

extern "C" int mtip_group_topk_prec(int precision, const void *workspace, int n_win, int64_t N_i,
                                    int64_t n_groups, const int32_t *child_off,
                                    const int32_t *child_idx, const void *slices, int64_t ld_s,
                                    const float *slice_scale, const double *slice_sumsq,
                                    const void *images, int64_t ld_d, const float *image_scale,
                                    int64_t kd, int K, int32_t *top_idx, float *top_val,
                                    int device, void *stream)
{
    return group_topk("mtip_group_topk_prec", precision, workspace, n_win, N_i, n_groups,
                      child_off, child_idx, slices, ld_s, slice_scale, slice_sumsq, images, ld_d,
                      image_scale, kd, K, top_idx, top_val, device, stream);
}

extern "C" int mtip_group_bad_count(const void *workspace, int n_win, int64_t N_i,
                                    int64_t n_groups, int32_t *out, void *stream)
{
    MTIP_REQUIRE(workspace && out && n_win >= 1, MTIP_ERR_ARG, "mtip_group_bad_count: bad args");
    GroupWs w = carve(const_cast<void *>(workspace), N_i * n_win, n_groups);
    MTIP_HIP_CHECK(hipMemcpyAsync(out, w.bad, sizeof(int32_t), hipMemcpyDeviceToDevice,
                                  (hipStream_t)stream));
    return MTIP_OK;
}

namespace {

// One wave per row: an entry equal to an earlier entry of its row becomes -1.
__global__ __launch_bounds__(256) void k_unique_rows(int32_t *__restrict__ idx, int64_t rows,
                                                     int cols)
{
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= rows) return;
    int32_t *r = idx + row * cols;
    for (int c0 = 0; c0 < cols; c0 += 64) {
        const int c = c0 + lane;
        const int32_t v = c < cols ? r[c] : -1;
        bool dup = false;
        for (int j = 0; j < c && !dup; ++j) dup = r[j] == v;
        if (c < cols && dup) r[c] = -1;
    }
}

}  // namespace

extern "C" int mtip_unique_rows(int32_t *idx, int64_t rows, int cols, void *stream)
{
    MTIP_REQUIRE(rows >= 0 && cols >= 1, MTIP_ERR_ARG, "mtip_unique_rows: bad sizes");
    if (rows == 0) return MTIP_OK;
    MTIP_REQUIRE(idx, MTIP_ERR_ARG, "mtip_unique_rows: null pointer");
    hipLaunchKernelGGL(k_unique_rows, dim3((unsigned)mtip::ceil_div(rows, 4)), dim3(256), 0,
                       (hipStream_t)stream, idx, rows, cols);
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}

extern "C" int mtip_merge_distinct(const int32_t *in_idx, const float *in_val, int64_t N_i,
                                   int n_in, int K_out, int32_t *out_idx, float *out_val,
                                   void *stream)
{
    MTIP_REQUIRE(N_i >= 0 && n_in >= 1, MTIP_ERR_ARG, "mtip_merge_distinct: bad sizes");
    MTIP_REQUIRE(K_out == 1 || K_out == 2 || K_out == 4 || K_out == 8, MTIP_ERR_ARG,
                 "mtip_merge_distinct: K_out must be 1, 2, 4 or 8");
    if (N_i == 0) return MTIP_OK;
    MTIP_REQUIRE(in_idx && in_val && out_idx && out_val, MTIP_ERR_ARG,
                 "mtip_merge_distinct: null pointer");
    hipStream_t st = (hipStream_t)stream;
    const dim3 grid((unsigned)mtip::ceil_div(N_i, 256));
    switch (K_out) {
    case 1: hipLaunchKernelGGL(k_merge_distinct<1>, grid, dim3(256), 0, st, in_idx, in_val, N_i, n_in, out_idx, out_val); break;
    case 2: hipLaunchKernelGGL(k_merge_distinct<2>, grid, dim3(256), 0, st, in_idx, in_val, N_i, n_in, out_idx, out_val); break;
    case 4: hipLaunchKernelGGL(k_merge_distinct<4>, grid, dim3(256), 0, st, in_idx, in_val, N_i, n_in, out_idx, out_val); break;
    default: hipLaunchKernelGGL(k_merge_distinct<8>, grid, dim3(256), 0, st, in_idx, in_val, N_i, n_in, out_idx, out_val); break;
    }
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}
