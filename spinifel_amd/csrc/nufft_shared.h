// Pieces of the ES-kernel NUFFT shared by the type-2 (nufft.hip) and
// type-1 (adjoint.hip) translation units: the plan, the grid view, the
// kernel function, window geometry, lane-group helpers and launch helpers.
#pragma once

#include <math.h>
#include <type_traits>

#include "mtip_internal.h"

struct mtip_plan {
    int device = 0;
    int M = 0, h = 0, K = 0;      // model edge, half width, symmetric extent 2h+1
    int n = 0, w = 0, pad = 0, np = 0, off = 0;
    double beta = 0.0, eps = 0.0;
    float *grid = nullptr;        // (np)^3 real fine grid
    double *dinv = nullptr;       // K: 1 / phihat(k)
    double2 *tw = nullptr;        // n: exp(-2 pi i m / n)
    double2 *G = nullptr;         // n^3: the deconvolved model at k mod n, then its FFT
    void *fft = nullptr;          // hipfftHandle of the n^3 Z2Z transform
    bool loaded = false;
    // type-1 (adjoint) state, allocated on first use
    double *acc = nullptr;        // (np)^3 FP64 spreading grid
    double2 *a1 = nullptr;        // np * np * (h + 1)
    double2 *a2 = nullptr;        // np * K * (h + 1)
    bool spread_open = false;     // mtip_plan_spread_begin called
};

namespace {

typedef unsigned int uint4v __attribute__((ext_vector_type(4)));
typedef unsigned int uint2v __attribute__((ext_vector_type(2)));

enum { GEOM_ROTATE = 0, GEOM_POINTS = 1 };

struct GridView {
    const float *__restrict__ g;
    int np, off, n;
    uint32_t bytes;       // size of the padded grid in bytes (buffer range)
    float beta_l2e, c2;   // beta * log2(e), 4 / w^2
    double inv_n;
};

// phi(z) = exp(beta (sqrt(1 - c2 z^2) - 1)) for c2 z^2 < 1, else 0.
__device__ __forceinline__ float es_phi(float z, float beta_l2e, float c2)
{
    // hardware sqrt / exp2 (about 1 ulp): far below the kernel's own error
    float v = 1.0f - c2 * z * z;
    return v > 0.0f ? __builtin_amdgcn_exp2f(beta_l2e * (__builtin_amdgcn_sqrtf(v) - 1.0f)) : 0.0f;
}


__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Inverse of skopi's rotation of quaternion [w, x, y, z] (eval/geom.py:26-46),
// by the adjugate in FP64 (np.linalg.inv at orientation_matching.py:220).
__device__ __forceinline__ void quat_inv_rot(const double *q4, double (&ri)[3][3])
{
    double w = q4[0], x = q4[1], y = q4[2], z = q4[3];
    double q01 = w * x, q02 = w * y, q03 = w * z, q11 = x * x, q12 = x * y, q13 = x * z;
    double q22 = y * y, q23 = y * z, q33 = z * z;
    double r[3][3];
    r[0][0] = 1.0 - 2.0 * (q22 + q33);
    r[0][1] = 2.0 * (q12 - q03);
    r[0][2] = 2.0 * (q13 + q02);
    r[1][0] = 2.0 * (q12 + q03);
    r[1][1] = 1.0 - 2.0 * (q11 + q33);
    r[1][2] = 2.0 * (q23 - q01);
    r[2][0] = 2.0 * (q13 - q02);
    r[2][1] = 2.0 * (q23 + q01);
    r[2][2] = 1.0 - 2.0 * (q11 + q22);
    double c00 = r[1][1] * r[2][2] - r[1][2] * r[2][1];
    double c01 = r[1][2] * r[2][0] - r[1][0] * r[2][2];
    double c02 = r[1][0] * r[2][1] - r[1][1] * r[2][0];
    double det = r[0][0] * c00 + r[0][1] * c01 + r[0][2] * c02;
    double id = 1.0 / det;
    ri[0][0] = c00 * id;
    ri[1][0] = c01 * id;
    ri[2][0] = c02 * id;
    ri[0][1] = (r[0][2] * r[2][1] - r[0][1] * r[2][2]) * id;
    ri[1][1] = (r[0][0] * r[2][2] - r[0][2] * r[2][0]) * id;
    ri[2][1] = (r[0][1] * r[2][0] - r[0][0] * r[2][1]) * id;
    ri[0][2] = (r[0][1] * r[1][2] - r[0][2] * r[1][1]) * id;
    ri[1][2] = (r[0][2] * r[1][0] - r[0][0] * r[1][2]) * id;
    ri[2][2] = (r[0][0] * r[1][1] - r[0][1] * r[1][0]) * id;
}

__device__ __forceinline__ double block_sum_f64(double v, double *s_red)
{
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) s_red[wid] = v;
    __syncthreads();
    double t = 0.0;
    int nw = blockDim.x >> 6;
    for (int i = 0; i < nw; ++i) t += s_red[i];
    return t;
}

// 47 KiB of box: with the point arrays a workgroup takes under 53.4 KiB of
// LDS, so three fit the 160 KiB of a CU.
constexpr int kBoxCap = 12032;
// 16-byte staging chunks per thread loaded in one batch: a typical box at
// M = 81 needs about 8, the largest within kBoxCap up to 15 (the rest follow
// in batches of 4; a larger first batch costs registers and occupancy)
constexpr int kStageChunks = 8;

template <int W>
struct PairTaps {
    static constexpr int NP = (W + 2) / 2;                                  // pairs per window
    static constexpr int GP = NP <= 2 ? 2 : NP <= 4 ? 4 : NP <= 8 ? 8 : 16;  // lanes per point
    static constexpr int NS = (W + GP - 1) / GP;                            // weights per lane
};

// Sum over aligned groups of G lanes; the result is valid in lane 0 of each
// group.  DPP moves (VALU) instead of LDS permutes: quad_perm for xor 1 / 2,
// row_shl 4 / 8 for the other quads of the 8- / 16-lane group.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float x)
{
    // mov_dpp has no "old" operand: lanes whose source is outside the row are
    // left undefined instead of taking a value that first has to be written
    // (update_dpp(0, ...) costs one v_mov per use).  Callers read only lanes
    // with a valid source: quad_perm sources are always valid, and after
    // row_shl the group sums are read in lane 0 of each group.
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xf, 0xf, false));
}

template <int G>
__device__ __forceinline__ float group_sum(float v)
{
    if (G >= 2) v += dpp_f<0xB1>(v);     // quad_perm [1,0,3,2]: lane ^ 1
    if (G >= 4) v += dpp_f<0x4E>(v);     // quad_perm [2,3,0,1]: lane ^ 2
    if (G >= 8) v += dpp_f<0x104>(v);    // row_shl:4: lane + 4 within the row of 16
    if (G >= 16) v += dpp_f<0x108>(v);   // row_shl:8: lane + 8 within the row of 16
    return v;
}

// Value of lane K of this lane's aligned group of GP lanes.
template <int GP, int K>
__device__ __forceinline__ float group_bcast(float x, int lane)
{
    if constexpr (GP == 2)
        return dpp_f<K | (K << 2) | ((K + 2) << 4) | ((K + 2) << 6)>(x);   // quad_perm [K,K,K+2,K+2]
    else if constexpr (GP == 4)
        return dpp_f<0x55 * K>(x);                                         // quad_perm [K,K,K,K]
    else
        return __shfl(x, (lane & ~(GP - 1)) + K, 64);
}

// w[a] = weight a of the group's point; lane c holds weights c, c + GP, ...
template <int GP, int W, int A = 0>
__device__ __forceinline__ void gather_weights(const float (&own)[PairTaps<W>::NS], float (&w)[W],
                                               int lane)
{
    if constexpr (A < W) {
        w[A] = group_bcast<GP, A % GP>(own[A / GP], lane);
        gather_weights<GP, W, A + 1>(own, w, lane);
    }
}

// Minimum over the wave, in every lane: DPP within each row of 16 lanes
// (VALU; the xor shuffles this replaces were LDS permutes, and the LDS is the
// slicing kernel's binding unit), then the four row results by readlane.
__device__ __forceinline__ int wave_min_i(int v)
{
    v = min(v, __builtin_amdgcn_mov_dpp(v, 0xB1, 0xf, 0xf, false));    // quad_perm [1,0,3,2]
    v = min(v, __builtin_amdgcn_mov_dpp(v, 0x4E, 0xf, 0xf, false));    // quad_perm [2,3,0,1]
    v = min(v, __builtin_amdgcn_mov_dpp(v, 0x124, 0xf, 0xf, false));   // row_ror:4
    v = min(v, __builtin_amdgcn_mov_dpp(v, 0x128, 0xf, 0xf, false));   // row_ror:8
    const int r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    const int r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    return min(min(r0, r1), min(r2, r3));
}

// Window corner as separate integer coordinates (padded-grid indices).
template <int W>
__device__ __forceinline__ bool point_corner(const GridView &gv, double X0, double X1, double X2,
                                             int (&ijk)[3], float (&t)[3])
{
    const double lim = 1.5 * gv.n;
    if (!(fabs(X0) < lim && fabs(X1) < lim && fabs(X2) < lim)) return false;
    const double dn = (double)gv.n;
    double X[3] = {X0, X1, X2};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double x = X[d] - dn * floor(X[d] * gv.inv_n + 0.5);
        int s = (int)ceil(x - 0.5 * W);
        s = max(s, -gv.n / 2 - W);
        s = min(s, gv.n / 2 + 1);
        ijk[d] = s + gv.off;
        t[d] = (float)(x - (double)s);
    }
    return true;
}

// Even LDS row strides the tap loop is compiled for.
#define MTIP_PAIR_STRIDES(X)                                                                  \
    X(8) X(10) X(12) X(14) X(16) X(18) X(20) X(22) X(24) X(26) X(28) X(30) X(32) X(36) X(40)  \
    X(48) X(56) X(64)

__device__ __forceinline__ int pair_stride(int bz)
{
    bz += bz & 1;
    return bz <= 8 ? 8 : bz <= 32 ? bz : bz <= 36 ? 36 : bz <= 40 ? 40 : bz <= 48 ? 48
         : bz <= 56 ? 56 : bz <= 64 ? 64 : 0;
}

#define MTIP_W_DISPATCH(W, FN, ...)                                                          \
    switch (W) {                                                                             \
    case 2: FN<2>(__VA_ARGS__); break;                                                       \
    case 3: FN<3>(__VA_ARGS__); break;                                                       \
    case 4: FN<4>(__VA_ARGS__); break;                                                       \
    case 5: FN<5>(__VA_ARGS__); break;                                                       \
    case 6: FN<6>(__VA_ARGS__); break;                                                       \
    case 7: FN<7>(__VA_ARGS__); break;                                                       \
    case 8: FN<8>(__VA_ARGS__); break;                                                       \
    case 9: FN<9>(__VA_ARGS__); break;                                                       \
    case 10: FN<10>(__VA_ARGS__); break;                                                     \
    case 11: FN<11>(__VA_ARGS__); break;                                                     \
    case 12: FN<12>(__VA_ARGS__); break;                                                     \
    case 13: FN<13>(__VA_ARGS__); break;                                                     \
    case 14: FN<14>(__VA_ARGS__); break;                                                     \
    case 15: FN<15>(__VA_ARGS__); break;                                                     \
    case 16: FN<16>(__VA_ARGS__); break;                                                     \
    default: return mtip::fail(MTIP_ERR_UNSUPPORTED, "kernel width out of range");          \
    }

GridView view_of(const mtip_plan *p)
{
    GridView gv;
    gv.g = p->grid;
    gv.np = p->np;
    gv.off = p->off;
    gv.n = p->n;
    gv.beta_l2e = (float)(p->beta * 1.4426950408889634);
    gv.c2 = 4.0f / (float)(p->w * p->w);
    gv.bytes = (uint32_t)((size_t)p->np * p->np * p->np * sizeof(float));
    gv.inv_n = 1.0 / (double)p->n;
    return gv;
}

struct DeviceGuard {
    int prev = 0;
    bool ok = false;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) == hipSuccess && hipSetDevice(dev) == hipSuccess) ok = true;
    }
    ~DeviceGuard()
    {
        if (ok) (void)hipSetDevice(prev);
    }
};

}  // namespace
