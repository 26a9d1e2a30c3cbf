// Type-2 3D NUFFT of a real cubic model on gfx950: plan, model load (fine
// grid) and the fused rotation + interpolation slice kernel.
//
// Method (exponential-of-semicircle kernel, upsampling sigma = 2):
//   phi(z) = exp(beta (sqrt(1 - (2 z / w)^2) - 1)),  |z| < w / 2
//   b_l    = sum_k g_k exp(-2 pi i k . l / n),  g_k = sym(a)_k / prod phihat(k_d)
//   Re c_j = sum_l b_l prod_d phi(X_jd - l_d),  X = x n / (2 pi)
// sym(a)_k = (a(k) + a(-k)) / 2 makes b real (the reference keeps only the
// real part of the transform); b comes from one FP64 hipFFT transform of the
// n^3 grid, and the fine grid is a real float32 array of
// (n + 2 pad)^3 values, indices l in [-n/2 - pad, n/2 + pad), evaluated
// directly at the padded indices so no wrap is needed while interpolating.
#include <math.h>
#include <vector>

#include <hipfft/hipfft.h>

#include "nufft_shared.h"

#include <type_traits>

namespace {

constexpr int kMaxW = 16;

int next235even(int n)
{
    if (n <= 2) return 2;
    if (n % 2) ++n;
    for (;; n += 2) {
        int m = n;
        for (int p : {2, 3, 5})
            while (m % p == 0) m /= p;
        if (m == 1) return n;
    }
}

// Gauss-Legendre nodes/weights on [-1, 1] (Newton on the three-term recurrence).
void gauss_legendre(int q, std::vector<double> &x, std::vector<double> &wts)
{
    x.assign(q, 0.0);
    wts.assign(q, 0.0);
    for (int i = 0; i < (q + 1) / 2; ++i) {
        double z = cos(M_PI * (i + 0.75) / (q + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= q; ++k) {
                double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            double pn = (q == 1) ? z : p1, pm = (q == 1) ? 1.0 : p0;
            double dp = q * (z * pn - pm) / (z * z - 1.0);
            double dz = pn / dp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        double p0 = 1.0, p1 = z;
        for (int k = 2; k <= q; ++k) {
            double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        double dp = q * (z * p1 - p0) / (z * z - 1.0);
        x[i] = -z;
        x[q - 1 - i] = z;
        wts[i] = wts[q - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

double es_phi_host(double z, int w, double beta)
{
    double t = 2.0 * z / w, v = 1.0 - t * t;
    return v > 0.0 ? exp(beta * (sqrt(v) - 1.0)) : 0.0;
}

// ------------------------------------------------------------ fine grid

__global__ void k_twiddle(double2 *tw, int n)
{
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < n) {
        double s, c;
        sincospi(2.0 * (double)m / (double)n, &s, &c);
        tw[m] = make_double2(c, -s);
    }
}

// Fine grid by FFT: G holds g_k = sym(a)_k / prod phihat(k_d) at index
// k mod n (zero elsewhere), one thread per G entry so no clearing pass is
// needed; hipFFT's forward Z2Z gives b_l = sum_k g_k exp(-2 pi i k . l / n),
// real because g is real and even; k_pad_grid writes b at the padded indices
// (periodic in l) as float32.
__global__ __launch_bounds__(256) void k_scatter_model(const float *__restrict__ ac,
                                                       const double *__restrict__ dinv,
                                                       double2 *__restrict__ G, int M, int h,
                                                       int n)
{
    const int64_t n3 = (int64_t)n * n * n;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    const int i3 = (int)(i % n), i2 = (int)((i / n) % n), i1 = (int)(i / ((int64_t)n * n));
    const int k1 = i1 <= n / 2 ? i1 : i1 - n, k2 = i2 <= n / 2 ? i2 : i2 - n,
              k3 = i3 <= n / 2 ? i3 : i3 - n;
    double v = 0.0;
    if (abs(k1) <= h && abs(k2) <= h && abs(k3) <= h) {
        const int lo = -(M / 2), hi = (M - 1) / 2;
        double s = 0.0;
        if (k1 >= lo && k1 <= hi && k2 >= lo && k2 <= hi && k3 >= lo && k3 <= hi)
            s += ac[((int64_t)(k1 + M / 2) * M + (k2 + M / 2)) * M + (k3 + M / 2)];
        if (-k1 >= lo && -k1 <= hi && -k2 >= lo && -k2 <= hi && -k3 >= lo && -k3 <= hi)
            s += ac[((int64_t)(-k1 + M / 2) * M + (-k2 + M / 2)) * M + (-k3 + M / 2)];
        v = 0.5 * s * dinv[k1 + h] * dinv[k2 + h] * dinv[k3 + h];
    }
    G[i] = make_double2(v, 0.0);
}

__global__ __launch_bounds__(256) void k_pad_grid(const double2 *__restrict__ G,
                                                  float *__restrict__ grid, int n, int np, int off)
{
    const int64_t np3 = (int64_t)np * np * np;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np3) return;
    const int p3 = (int)(i % np), p2 = (int)((i / np) % np), p1 = (int)(i / ((int64_t)np * np));
    auto wrap = [n, off](int p) { return ((p - off) % n + n) % n; };
    grid[i] = (float)G[((int64_t)wrap(p1) * n + wrap(p2)) * n + wrap(p3)].x;
}

// ------------------------------------------------------------ interpolation
//
// Lane layout: a group of G lanes (G = 4, 8 or 16 >= w) interpolates one
// point; lane c of the group owns tap c along the grid's fast axis (L), so
// every row read of the w x w x w window is one contiguous G-float segment
// shared by the group, fetched with a buffer load whose row offset
// (a * np^2 + b * np) is a wave-uniform SGPR.  Per 64-point chunk, each lane
// first computes the geometry of one point (FP64 rotation, window start,
// FP32 fractions) into LDS; the wave then walks the chunk G points... 64/G
// points at a time.  Lane c evaluates the three kernel weights at offset c;
// the group shares them through LDS; the group sums its lanes' partial
// results with xor-shuffles.

// Window corner (flat element index into the padded grid) and fractional
// offsets of a point at grid coordinates X (grid units).  Returns false for
// points outside |x| < 3 pi (|X| < 1.5 n) or not finite.  The coordinate is
// wrapped into [-n/2, n/2) and the window start clamped so that every access
// stays inside the padded grid whatever the rounding.
template <int W>
__device__ __forceinline__ bool point_geometry(const GridView &gv, double X0, double X1, double X2,
                                               int &base, float &t0, float &t1, float &t2)
{
    const double lim = 1.5 * gv.n;
    if (!(fabs(X0) < lim && fabs(X1) < lim && fabs(X2) < lim)) return false;
    const double dn = (double)gv.n;
    double X[3] = {X0, X1, X2};
    int i[3];
    float t[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double x = X[d] - dn * floor(X[d] * gv.inv_n + 0.5);
        int s = (int)ceil(x - 0.5 * W);
        s = max(s, -gv.n / 2 - W);
        s = min(s, gv.n / 2 + 1);
        i[d] = s + gv.off;
        t[d] = (float)(x - (double)s);
    }
    base = (i[0] * gv.np + i[1]) * gv.np + i[2];
    t0 = t[0];
    t1 = t[1];
    t2 = t[2];
    return true;
}

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// GEOM_ROTATE: one workgroup per orientation (blockIdx.x), points are the
// N_p pixel q-vectors q (3, N_p) rotated by inv(R(quat)) and scaled by
// gscale into grid units; out row o, plus the row's FP64 sum / sum of squares.
// GEOM_POINTS: 256 points per workgroup from x, y, z (radians), gscale = n / 2pi.
template <int W, int G, int GEOM>
__global__ __launch_bounds__(256) void k_interp(GridView gv, const double *__restrict__ quats,
                                                const double *__restrict__ qx,
                                                const double *__restrict__ qy,
                                                const double *__restrict__ qz, int64_t N_p,
                                                double gscale, float *__restrict__ out,
                                                int64_t ld_out, uint16_t *__restrict__ out_bf16,
                                                double *__restrict__ row_sum,
                                                double *__restrict__ row_sumsq, int *bad)
{
    constexpr int PPW = 64 / G;    // points per wave per step
    __shared__ int s_base[4][64];
    __shared__ float4 s_t[4][64];
    __shared__ float s_w[4][PPW][2][G];
    __shared__ double s_red[8];

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane / G, c = lane % G;
    const int64_t o = (GEOM == GEOM_ROTATE) ? (int64_t)blockIdx.x : 0;
    double ri[3][3];
    if (GEOM == GEOM_ROTATE) quat_inv_rot(quats + 4 * o, ri);
    const int64_t p0 = (GEOM == GEOM_ROTATE) ? 0 : (int64_t)blockIdx.x * 256;
    const int64_t p1 = (GEOM == GEOM_ROTATE) ? N_p : (p0 + 256 < N_p ? p0 + 256 : N_p);
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc((void *)gv.g, (short)0, (int)gv.bytes, 0x00020000);
    const int np = gv.np;
    float *orow = out + o * ld_out;
    uint16_t *brow = out_bf16 ? out_bf16 + o * ld_out : nullptr;
    double sum = 0.0, sumsq = 0.0;
    int nbad = 0;

    for (int64_t cb = p0 + wave * 64; cb < p1; cb += 256) {
        // geometry of point cb + lane
        {
            const int64_t j = cb + lane;
            int base = -1;
            float t0 = 0.f, t1 = 0.f, t2 = 0.f;
            if (j < p1) {
                double X0, X1, X2;
                if (GEOM == GEOM_ROTATE) {
                    double x = qx[j], y = qx[N_p + j], z = qx[2 * N_p + j];
                    X0 = (ri[0][0] * x + ri[0][1] * y + ri[0][2] * z) * gscale;
                    X1 = (ri[1][0] * x + ri[1][1] * y + ri[1][2] * z) * gscale;
                    X2 = (ri[2][0] * x + ri[2][1] * y + ri[2][2] * z) * gscale;
                } else {
                    X0 = qx[j] * gscale;
                    X1 = qy[j] * gscale;
                    X2 = qz[j] * gscale;
                }
                if (!point_geometry<W>(gv, X0, X1, X2, base, t0, t1, t2)) base = -2;
            }
            s_base[wave][lane] = base;
            s_t[wave][lane] = make_float4(t0, t1, t2, 0.f);
        }
        wave_sync();
#pragma unroll 1
        for (int it = 0; it < G; ++it) {
            const int pl = it * PPW + g;
            const int pb = s_base[wave][pl];
            const float4 tt = s_t[wave][pl];
            float wa_c = 0.f, wb_c = 0.f, wc_c = 0.f;
            if (c < W) {
                wa_c = es_phi((float)c - tt.x, gv.beta_l2e, gv.c2);
                wb_c = es_phi((float)c - tt.y, gv.beta_l2e, gv.c2);
                wc_c = es_phi((float)c - tt.z, gv.beta_l2e, gv.c2);
            }
            s_w[wave][g][0][c] = wa_c;
            s_w[wave][g][1][c] = wb_c;
            wave_sync();
            float wa[W], wb[W];
#pragma unroll
            for (int a = 0; a < W; ++a) {
                wa[a] = s_w[wave][g][0][a];
                wb[a] = s_w[wave][g][1][a];
            }
            const int voff = ((pb >= 0 ? pb : 0) + c) * 4;
            // issue all W*W row loads first so their latencies overlap
            uint32_t raw[W][W];
#pragma unroll
            for (int a = 0; a < W; ++a)
#pragma unroll
                for (int b = 0; b < W; ++b)
                    raw[a][b] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, voff,
                                                                     (a * np * np + b * np) * 4, 0);
            float acc = 0.f;
#pragma unroll
            for (int a = 0; a < W; ++a) {
                float tab = 0.f;
#pragma unroll
                for (int b = 0; b < W; ++b) tab = fmaf(wb[b], __uint_as_float(raw[a][b]), tab);
                acc = fmaf(wa[a], tab, acc);
            }
            float r = acc * wc_c;
#pragma unroll
            for (int m = G / 2; m >= 1; m >>= 1) r += __shfl_xor(r, m, 64);
            if (c == 0 && pb != -1) {
                const int64_t j = cb + pl;
                float v = r;
                if (pb == -2) {
                    v = __builtin_nanf("");
                    ++nbad;
                }
                orow[j] = v;
                if (brow) brow[j] = f32_to_bf16_rne(v);
                sum += (double)v;
                sumsq += (double)v * (double)v;
            }
            wave_sync();
        }
    }
    if (nbad && bad) atomicAdd(bad, nbad);
    if (GEOM == GEOM_ROTATE && (row_sum || row_sumsq)) {
        double ts = block_sum_f64(sum, s_red);
        double tq = block_sum_f64(sumsq, s_red);
        if (threadIdx.x == 0) {
            if (row_sum) row_sum[o] = ts;
            if (row_sumsq) row_sumsq[o] = tq;
        }
    }
}

// ------------------------------------------------------------ LDS-staged slicing
//
// One workgroup per orientation walks the detector in tiles of TP = 256, 128
// or 64 pixels (16 x 16, 8 x 16 or 8 x 8 blocks when the caller passes a tile
// order; the smaller tiles keep the boxes of large grids within LDS).  A tile's points lie
// on a small patch of the rotated Ewald sphere, so the union of their w^3
// windows fits a small box of the fine grid (about 8k floats at M = 81, 0.9
// Nyquist): the workgroup stages that box in LDS with 16-byte loads, then
// runs the taps from LDS.  GP lanes serve one point: lane c owns the pair of
// fast-axis elements (2c, 2c + 1) counted from the even element at or below
// the window start, so each LDS read is one aligned 8-byte read and elements
// outside the window get a zero weight.  The other two axes' weights move
// between the group's lanes by DPP broadcasts.  A tile whose box exceeds
// kBoxCap reads the grid directly (well under 1 % of tiles at M = 81).

// Taps of one lane's element pair from the staged box: rows b of plane a at
// compile-time offsets b * S, planes at runtime offsets a * plane.
constexpr int kPlanesInFlight = 3;

template <int W, int S>
__device__ __forceinline__ void taps_pair_lds(const float *__restrict__ box, int base, int plane,
                                              const float (&wa)[W], const float (&wb)[W],
                                              float &r0, float &r1)
{
    float a0 = 0.f, a1 = 0.f;
#pragma unroll
    for (int a = 0; a < W; ++a) {
        const float2 *pa = reinterpret_cast<const float2 *>(box + base + a * plane);
        float t0 = 0.f, t1 = 0.f;
#pragma unroll
        for (int b = 0; b < W; ++b) {
            const float2 v = pa[b * (S / 2)];
            t0 = fmaf(wb[b], v.x, t0);
            t1 = fmaf(wb[b], v.y, t1);
        }
        a0 = fmaf(wa[a], t0, a0);
        a1 = fmaf(wa[a], t1, a1);
        // issue the reads kPlanesInFlight planes at a time, each group's
        // reads ahead of the previous group's FMAs: the compiler otherwise
        // pairs and hoists nearly all W^2 reads above the first FMA (a
        // scheduling barrier does not stop the load pairing, a compiler
        // memory barrier does).  At M = 81, 10k x 16k: 8.02 ms with every read
        // hoisted, 8.31 ms one plane at a time, 7.87 ms three planes at a time
        if ((a + 1) % kPlanesInFlight == 0) asm volatile("" ::: "memory");
    }
    r0 = a0;
    r1 = a1;
}

// The same taps read from the padded grid (voff = byte offset of the pair).
template <int W>
__device__ __forceinline__ void taps_pair_direct(__amdgpu_buffer_rsrc_t rsrc, int np, int voff,
                                                 const float (&wa)[W], const float (&wb)[W],
                                                 float &r0, float &r1)
{
    // planes in two halves, each half's loads issued before its FMAs
    constexpr int HA = (W + 1) / 2;
    float a0 = 0.f, a1 = 0.f;
#pragma unroll
    for (int a_lo = 0; a_lo < W; a_lo += HA) {
        uint2v raw[HA][W];
#pragma unroll
        for (int i = 0; i < HA; ++i)
#pragma unroll
            for (int b = 0; b < W; ++b)
                if (a_lo + i < W)
                    raw[i][b] = __builtin_amdgcn_raw_buffer_load_b64(
                        rsrc, voff, ((a_lo + i) * np * np + b * np) * 4, 0);
#pragma unroll
        for (int i = 0; i < HA; ++i) {
            if (a_lo + i >= W) break;
            float t0 = 0.f, t1 = 0.f;
#pragma unroll
            for (int b = 0; b < W; ++b) {
                t0 = fmaf(wb[b], __uint_as_float(raw[i][b][0]), t0);
                t1 = fmaf(wb[b], __uint_as_float(raw[i][b][1]), t1);
            }
            a0 = fmaf(wa[a_lo + i], t0, a0);
            a1 = fmaf(wa[a_lo + i], t1, a1);
        }
    }
    r0 = a0;
    r1 = a1;
}

template <int W, int TP>
__global__ __launch_bounds__(256, 3) void k_slices_lds(GridView gv, const double *__restrict__ quats,
                                                    const double *__restrict__ q, int64_t N_p,
                                                    const int32_t *__restrict__ order,
                                                    int64_t n_order,
                                                    const int32_t *__restrict__ orient_order,
                                                    double gscale,
                                                    float *__restrict__ out, int64_t ld_out,
                                                    uint16_t *__restrict__ out_bf16,
                                                    double *__restrict__ row_sum,
                                                    double *__restrict__ row_sumsq, int *bad)
{
    using PT = PairTaps<W>;
    constexpr int GP = PT::GP, NS = PT::NS, PPW = 64 / GP;
    constexpr int PW = TP / 4;                                  // points per wave
    constexpr int STEPS = PW >= PPW ? PW / PPW : 1;
    __shared__ __align__(16) float s_box[kBoxCap];
    // per point: pair-base offset (box-local when staged, padded grid otherwise)
    // with the window parity in bit 0, or -1 for a point outside the domain
    __shared__ int s_lb[TP];
    __shared__ int s_pix[TP];          // output pixel, -1 for an empty slot
    __shared__ float s_t[3][TP];       // fractional window offsets
    __shared__ int s_red_i[4][6];
    __shared__ double s_red[8];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane / GP, c = lane % GP;
    // Workgroups go round-robin over the 8 XCDs; give each XCD a contiguous
    // run of the (caller-ordered) orientations so that the ones resident on
    // an XCD at the same time are neighbours and share its L2.
    const int64_t nb = gridDim.x, n8 = nb / 8, b = blockIdx.x;
    const int64_t slot = b < 8 * n8 ? (b % 8) * n8 + b / 8 : b;
    const int64_t o = orient_order ? (int64_t)orient_order[slot] : slot;
    // R^-1 of this orientation, kept in LDS: as registers its nine doubles
    // would be live across the whole kernel
    __shared__ double s_ri[9];
    if (tid == 0) {
        double ri[3][3];
        quat_inv_rot(quats + 4 * o, ri);
#pragma unroll
        for (int e = 0; e < 9; ++e) s_ri[e] = ri[e / 3][e % 3];
    }
    __syncthreads();
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc((void *)gv.g, (short)0, (int)gv.bytes, 0x00020000);
    const int np = gv.np;
    float *orow = out + o * ld_out;
    uint16_t *brow = out_bf16 ? out_bf16 + o * ld_out : nullptr;
    double sum = 0.0, sumsq = 0.0;
    int nbad = 0;
    const int64_t n_slots = order ? n_order : mtip::ceil_div(N_p, TP) * TP;
    const bool has_pt = tid < TP;      // threads holding a point of the tile

    // pixel index and q-vector of this thread's point, prefetched one tile ahead
    // (and the pixel index two tiles ahead) so their latency hides behind taps
    auto pixel_at = [&](int64_t k) -> int {
        if (!has_pt) return -1;
        if (order) return (k < n_order) ? order[k] : -1;
        return (k < N_p) ? (int)k : -1;
    };
    int pix_n = pixel_at(tid);
    double qx_n = 0.0, qy_n = 0.0, qz_n = 0.0;
    {
        const int pp = pix_n >= 0 ? pix_n : 0;
        qx_n = q[pp];
        qy_n = q[N_p + pp];
        qz_n = q[2 * N_p + pp];
    }
    int pix_nn = pixel_at(TP + tid);

    for (int64_t tb = 0; tb < n_slots; tb += TP) {
        // ---- geometry of this thread's point
        const int pix = pix_n;
        const double x = qx_n, y = qy_n, z = qz_n;
        pix_n = pix_nn;
        {
            const int pp = pix_n >= 0 ? pix_n : 0;
            qx_n = q[pp];
            qy_n = q[N_p + pp];
            qz_n = q[2 * N_p + pp];
        }
        pix_nn = pixel_at(tb + 2 * TP + tid);
        bool ok = false;
        int ijk[3] = {0, 0, 0};
        float t[3] = {0.f, 0.f, 0.f};
        if (pix >= 0) {
            double X0 = (s_ri[0] * x + s_ri[1] * y + s_ri[2] * z) * gscale;
            double X1 = (s_ri[3] * x + s_ri[4] * y + s_ri[5] * z) * gscale;
            double X2 = (s_ri[6] * x + s_ri[7] * y + s_ri[8] * z) * gscale;
            ok = point_corner<W>(gv, X0, X1, X2, ijk, t);
        }
        // ---- box of the tile's windows (block min / max)
        {
            int v[6] = {ok ? ijk[0] : 0x7fffffff, ok ? ijk[1] : 0x7fffffff, ok ? ijk[2] : 0x7fffffff,
                        ok ? -ijk[0] : 0x7fffffff, ok ? -ijk[1] : 0x7fffffff, ok ? -ijk[2] : 0x7fffffff};
#pragma unroll
            for (int e = 0; e < 6; ++e) v[e] = wave_min_i(v[e]);
            if (lane == 0)
#pragma unroll
                for (int e = 0; e < 6; ++e) s_red_i[wave][e] = v[e];
        }
        __syncthreads();
        int lo[3], hi[3];
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            lo[e] = min(min(s_red_i[0][e], s_red_i[1][e]), min(s_red_i[2][e], s_red_i[3][e]));
            hi[e] = -min(min(s_red_i[0][e + 3], s_red_i[1][e + 3]),
                         min(s_red_i[2][e + 3], s_red_i[3][e + 3]));
        }
        const bool any = lo[0] != 0x7fffffff;
        const int lz = lo[2] & ~1;      // even, so every pair base is 8-byte aligned
        const int bx = any ? hi[0] - lo[0] + W : 0, by = any ? hi[1] - lo[1] + W : 0;
        const int bz = any ? hi[2] - lz + 2 * GP : 0;
        const int S = pair_stride(bz);
        // plane stride: by * S raised to 10 mod 32 floats when the box still
        // fits.  A quarter-wave serves a 2 x 2 pixel group whose window
        // corners differ by about one cell per axis; with 32 banks, neighbours
        // one plane apart then read other banks.  Measured over residues at
        // M = 81, 10k x 16k: 8.75 ms at 10 mod 32 against 8.98 ms unpadded (a
        // quarter-wave bank model ranks the residues alike; 0 and 6 are worst)
        int plane = by * S;
        {
            const int padded = plane + ((10 - plane) & 31);
            if (bx * padded <= kBoxCap) plane = padded;
        }
        const bool staged = any && S > 0 && bx * plane <= kBoxCap;
        // this point's pair base (box-local when staged, padded grid otherwise)
        int lb = -1;
        if (ok) {
            const int zb = ijk[2] & ~1;
            lb = staged ? (ijk[0] - lo[0]) * plane + (ijk[1] - lo[1]) * S + (zb - lz)
                        : (ijk[0] * np + ijk[1]) * np + zb;
            lb |= ijk[2] & 1;
        }

        if (has_pt) {
            s_lb[tid] = lb;
            s_pix[tid] = pix;
            s_t[0][tid] = t[0];
            s_t[1][tid] = t[1];
            s_t[2][tid] = t[2];
        }
        // ---- the box
        if (staged) {
            // the box as nrows * q4 chunks of 16 bytes, chunk k = row k / q4,
            // quarter k % q4: every thread's loads go in flight at once (at most
            // kStageChunks each, enough for any box within kBoxCap), then the
            // 8-byte LDS stores (pairs at or past S dropped)
            const int q4 = (bz + 3) >> 2;
            const int nchunk = bx * by * q4;
            const float inv_q4 = 1.0f / (float)q4, inv_by = 1.0f / (float)by;
            auto stage = [&](int u0, auto nb) __attribute__((always_inline)) {
                constexpr int NB = decltype(nb)::value;
                uint4v v[NB];
#pragma unroll
                for (int u = 0; u < NB; ++u) {
                    const int k = (u0 + u) * 256 + tid;
                    // k / q4 and r / by for k < 2^16: the float estimates are
                    // within half a unit of the integer quotient
                    const int r = (int)(((float)k + 0.5f) * inv_q4), zq = k - r * q4;
                    const int xr = (int)(((float)r + 0.5f) * inv_by), yr = r - xr * by;
                    const int gidx =
                        k < nchunk ? ((lo[0] + xr) * np + (lo[1] + yr)) * np + lz + 4 * zq : 0;
                    v[u] = __builtin_amdgcn_raw_buffer_load_b128(rsrc, gidx * 4, 0, 0);
                }
#pragma unroll
                for (int u = 0; u < NB; ++u) {
                    const int k = (u0 + u) * 256 + tid;
                    if (k < nchunk) {
                        const int r = (int)(((float)k + 0.5f) * inv_q4), z4 = 4 * (k - r * q4);
                        const int xr = (int)(((float)r + 0.5f) * inv_by), yr = r - xr * by;
                        float2 *d2 = reinterpret_cast<float2 *>(s_box + xr * plane + yr * S + z4);
                        d2[0] = make_float2(__uint_as_float(v[u][0]), __uint_as_float(v[u][1]));
                        if (z4 + 2 < S)
                            d2[1] = make_float2(__uint_as_float(v[u][2]), __uint_as_float(v[u][3]));
                    }
                }
            };
            stage(0, std::integral_constant<int, kStageChunks>{});
            for (int u0 = kStageChunks; u0 * 256 < nchunk; u0 += 4)
                stage(u0, std::integral_constant<int, 4>{});
        }
        __syncthreads();
        // ---- taps: each wave its PW points, PPW per step
#pragma unroll 1
        for (int it = 0; it < STEPS; ++it) {
            const bool slot_ok = PW >= PPW || g < PW;
            const int pl = slot_ok ? wave * PW + it * PPW + g : 0;
            const int lv = s_lb[pl];
            const float tx = s_t[0][pl], ty = s_t[1][pl], tz = s_t[2][pl];
            const int sh = lv & 1;
            const int base = lv >= 0 ? (lv & ~1) + 2 * c : 0;
            float own_a[NS], own_b[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                own_a[s] = es_phi((float)(c + GP * s) - tx, gv.beta_l2e, gv.c2);
                own_b[s] = es_phi((float)(c + GP * s) - ty, gv.beta_l2e, gv.c2);
            }
            float wa[W], wb[W];
            gather_weights<GP, W>(own_a, wa, lane);
            gather_weights<GP, W>(own_b, wb, lane);
            // elements 2c - sh and 2c + 1 - sh of the window (zero weight outside it)
            const float k0 = (float)(2 * c - sh);
            const float wc0 = es_phi(k0 - tz, gv.beta_l2e, gv.c2);
            const float wc1 = es_phi(k0 + 1.0f - tz, gv.beta_l2e, gv.c2);
            float r0 = 0.f, r1 = 0.f;
            if (staged) {
                switch (S) {
#define MTIP_TAPS_CASE(SS) case SS: taps_pair_lds<W, SS>(s_box, base, plane, wa, wb, r0, r1); break;
                    MTIP_PAIR_STRIDES(MTIP_TAPS_CASE)
#undef MTIP_TAPS_CASE
                default: break;
                }
            } else {
                taps_pair_direct<W>(rsrc, np, base * 4, wa, wb, r0, r1);
            }
            const float r = group_sum<GP>(fmaf(wc0, r0, wc1 * r1));
            // the slot's value replaces its x offset (read above, by this wave only)
            if (c == 0 && slot_ok) s_t[0][pl] = r;
        }
        // ---- outputs: each thread its own slot (written by its own wave when TP = 256)
        if constexpr (TP == 256) wave_sync();
        else __syncthreads();
        if (has_pt) {
            const int pix_o = s_pix[tid];
            if (pix_o >= 0) {
                float v = s_t[0][tid];
                if (s_lb[tid] < 0) {
                    v = __builtin_nanf("");
                    ++nbad;
                }
                orow[pix_o] = v;
                if (brow) brow[pix_o] = f32_to_bf16_rne(v);
                sum += (double)v;
                sumsq += (double)v * (double)v;
            }
        }
        __syncthreads();
    }
    if (nbad && bad) atomicAdd(bad, nbad);
    if (row_sum || row_sumsq) {
        double ts = block_sum_f64(sum, s_red);
        double tq = block_sum_f64(sumsq, s_red);
        if (threadIdx.x == 0) {
            if (row_sum) row_sum[o] = ts;
            if (row_sumsq) row_sumsq[o] = tq;
        }
    }
}

// Rotated, scaled q-points only (FP64): H/K/L[n * N_p + j] = scale * (R_n^-1 q_j).
__global__ __launch_bounds__(256) void k_positions(const double *__restrict__ quats,
                                                   const double *__restrict__ q, int64_t N_p,
                                                   double scale, double *__restrict__ H,
                                                   double *__restrict__ Kk, double *__restrict__ L)
{
    const int64_t o = blockIdx.y;
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N_p) return;
    double ri[3][3];
    quat_inv_rot(quats + 4 * o, ri);
    double qx = q[j], qy = q[N_p + j], qz = q[2 * N_p + j];
    const int64_t k = o * N_p + j;
    H[k] = (ri[0][0] * qx + ri[0][1] * qy + ri[0][2] * qz) * scale;
    Kk[k] = (ri[1][0] * qx + ri[1][1] * qy + ri[1][2] * qz) * scale;
    L[k] = (ri[2][0] * qx + ri[2][1] * qy + ri[2][2] * qz) * scale;
}

template <int W>
void launch_slices(const GridView &gv, const double *quats, int64_t N_o, const double *q,
                   int64_t N_p, const int32_t *order, int64_t n_order, int tile,
                   const int32_t *orient_order, double gscale, float *out, int64_t ld,
                   uint16_t *obf, double *rs, double *rq, int *bad, hipStream_t st)
{
#define MTIP_SLICES_TP(T)                                                                     \
    hipLaunchKernelGGL((k_slices_lds<W, T>), dim3((unsigned)N_o), dim3(256), 0, st, gv, quats, q, \
                       N_p, order, n_order, orient_order, gscale, out, ld, obf, rs, rq, bad)
    if (tile == 64) MTIP_SLICES_TP(64);
    else if (tile == 128) MTIP_SLICES_TP(128);
    else MTIP_SLICES_TP(256);
#undef MTIP_SLICES_TP
}

template <int W>
void launch_execute(const GridView &gv, const double *x, const double *y, const double *z,
                    int64_t N, float *out, int *bad, hipStream_t st)
{
    constexpr int G = W <= 4 ? 4 : (W <= 8 ? 8 : 16);
    hipLaunchKernelGGL((k_interp<W, G, GEOM_POINTS>), dim3((unsigned)mtip::ceil_div(N, 256)),
                       dim3(256), 0, st, gv, nullptr, x, y, z, N, (double)gv.n / (2.0 * M_PI),
                       out, (int64_t)0, nullptr, nullptr, nullptr, bad);
}

}  // namespace

extern "C" int mtip_plan_create(int M, double eps, int device, mtip_plan **out)
{
    MTIP_REQUIRE(out, MTIP_ERR_ARG, "mtip_plan_create: null output");
    *out = nullptr;
    MTIP_REQUIRE(M >= 2 && M <= 320, MTIP_ERR_ARG, "mtip_plan_create: M must be in [2, 320]");
    MTIP_REQUIRE(eps > 0.0 && eps < 1.0, MTIP_ERR_ARG, "mtip_plan_create: eps must be in (0, 1)");
    DeviceGuard dg(device);
    MTIP_REQUIRE(dg.ok, MTIP_ERR_HIP, "mtip_plan_create: cannot select device");
    mtip_plan *p = new mtip_plan();
    p->device = device;
    p->M = M;
    p->h = M / 2;
    p->K = 2 * p->h + 1;
    p->eps = eps;
    double e = eps < 1e-15 ? 1e-15 : eps;
    int w = (int)ceil(-log10(e / 10.0));
    w = w < 2 ? 2 : (w > kMaxW ? kMaxW : w);
    p->w = w;
    double bon = (w == 2) ? 2.20 : (w == 3) ? 2.26 : (w == 4) ? 2.38 : 2.30;
    p->beta = bon * w;
    int n0 = 2 * M > 2 * w ? 2 * M : 2 * w;
    p->n = next235even(n0);
    int G = w <= 4 ? 4 : (w <= 8 ? 8 : 16);
    p->pad = G + 1;     // window start clamp + G-lane row reads stay inside
    p->np = p->n + 2 * p->pad;
    p->off = p->n / 2 + p->pad;

    // 1 / phihat(k), k in [-h, h]: Gauss-Legendre on [-w/2, w/2].
    int nq = 2 * (2 + 3 * ((w + 1) / 2)) + 32;
    std::vector<double> gx, gw;
    gauss_legendre(nq, gx, gw);
    std::vector<double> dinv(p->K);
    for (int i = 0; i < p->K; ++i) {
        double k = i - p->h, s = 0.0;
        for (int t = 0; t < nq; ++t) {
            double xx = gx[t] * 0.5 * w;
            s += gw[t] * 0.5 * w * es_phi_host(xx, w, p->beta) * cos(2.0 * M_PI * k * xx / p->n);
        }
        dinv[i] = 1.0 / s;
    }
    size_t ng = (size_t)p->np * p->np * p->np;
    hipError_t err = hipMalloc(&p->grid, ng * sizeof(float));
    if (err == hipSuccess) err = hipMalloc(&p->dinv, p->K * sizeof(double));
    if (err == hipSuccess) err = hipMalloc(&p->tw, p->n * sizeof(double2));
    if (err == hipSuccess)
        err = hipMemcpy(p->dinv, dinv.data(), p->K * sizeof(double), hipMemcpyHostToDevice);
    if (err == hipSuccess) {
        hipLaunchKernelGGL(k_twiddle, dim3((p->n + 255) / 256), dim3(256), 0, 0, p->tw, p->n);
        err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipDeviceSynchronize();
    if (err != hipSuccess) {
        std::string msg = std::string("mtip_plan_create: ") + hipGetErrorString(err);
        mtip_plan_destroy(p);
        return mtip::fail(err == hipErrorOutOfMemory ? MTIP_ERR_NOMEM : MTIP_ERR_HIP, msg);
    }

    *out = p;
    return MTIP_OK;
}

extern "C" int mtip_plan_destroy(mtip_plan *p)
{
    if (!p) return MTIP_OK;
    DeviceGuard dg(p->device);
    (void)hipFree(p->grid);
    (void)hipFree(p->dinv);
    (void)hipFree(p->tw);
    (void)hipFree(p->G);
    if (p->fft) (void)hipfftDestroy((hipfftHandle)p->fft);
    (void)hipFree(p->acc);
    (void)hipFree(p->a1);
    (void)hipFree(p->a2);
    delete p;
    return MTIP_OK;
}

extern "C" int mtip_plan_info(const mtip_plan *p, int *n_fine, int *width, double *beta,
                              int *padded_edge)
{
    MTIP_REQUIRE(p, MTIP_ERR_ARG, "mtip_plan_info: null plan");
    if (n_fine) *n_fine = p->n;
    if (width) *width = p->w;
    if (beta) *beta = p->beta;
    if (padded_edge) *padded_edge = p->np;
    return MTIP_OK;
}

extern "C" int mtip_plan_set_model(mtip_plan *p, const float *ac, void *stream)
{
    MTIP_REQUIRE(p && ac, MTIP_ERR_ARG, "mtip_plan_set_model: null argument");
    DeviceGuard dg(p->device);
    MTIP_REQUIRE(dg.ok, MTIP_ERR_HIP, "mtip_plan_set_model: cannot select device");
    hipStream_t st = (hipStream_t)stream;
    const int64_t n3 = (int64_t)p->n * p->n * p->n;
    const int64_t np3 = (int64_t)p->np * p->np * p->np;
    if (!p->G) {
        // first model on this plan: the FFT buffer and plan (plans used only
        // for the adjoint never need them)
        hipError_t e = hipMalloc(&p->G, (size_t)n3 * sizeof(double2));
        if (e != hipSuccess) {
            p->G = nullptr;
            return mtip::fail(e == hipErrorOutOfMemory ? MTIP_ERR_NOMEM : MTIP_ERR_HIP,
                              std::string("mtip_plan_set_model: ") + hipGetErrorString(e));
        }
        hipfftHandle fh = nullptr;
        if (hipfftPlan3d(&fh, p->n, p->n, p->n, HIPFFT_Z2Z) != HIPFFT_SUCCESS) {
            (void)hipFree(p->G);
            p->G = nullptr;
            return mtip::fail(MTIP_ERR_HIP, "mtip_plan_set_model: hipfftPlan3d failed");
        }
        p->fft = fh;
    }
    hipLaunchKernelGGL(k_scatter_model, dim3((unsigned)mtip::ceil_div(n3, 256)), dim3(256), 0, st,
                       ac, p->dinv, p->G, p->M, p->h, p->n);
    MTIP_HIP_CHECK(hipGetLastError());
    hipfftHandle fh = (hipfftHandle)p->fft;
    MTIP_REQUIRE(hipfftSetStream(fh, st) == HIPFFT_SUCCESS, MTIP_ERR_HIP,
                 "mtip_plan_set_model: hipfftSetStream failed");
    MTIP_REQUIRE(hipfftExecZ2Z(fh, (hipfftDoubleComplex *)p->G, (hipfftDoubleComplex *)p->G,
                               HIPFFT_FORWARD) == HIPFFT_SUCCESS,
                 MTIP_ERR_HIP, "mtip_plan_set_model: hipfftExecZ2Z failed");
    hipLaunchKernelGGL(k_pad_grid, dim3((unsigned)mtip::ceil_div(np3, 256)), dim3(256), 0, st,
                       p->G, p->grid, p->n, p->np, p->off);
    MTIP_HIP_CHECK(hipGetLastError());
    p->loaded = true;
    return MTIP_OK;
}

extern "C" int mtip_plan_execute(mtip_plan *p, const double *x, const double *y, const double *z,
                                 int64_t N, float *out, int *bad_count, void *stream)
{
    MTIP_REQUIRE(p, MTIP_ERR_ARG, "mtip_plan_execute: null plan");
    MTIP_REQUIRE(p->loaded, MTIP_ERR_STATE, "mtip_plan_execute: no model set");
    MTIP_REQUIRE(N >= 0, MTIP_ERR_ARG, "mtip_plan_execute: N < 0");
    if (N == 0) return MTIP_OK;
    MTIP_REQUIRE(x && y && z && out, MTIP_ERR_ARG, "mtip_plan_execute: null pointer");
    DeviceGuard dg(p->device);
    GridView gv = view_of(p);
    MTIP_W_DISPATCH(p->w, launch_execute, gv, x, y, z, N, out, bad_count, (hipStream_t)stream);
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}

extern "C" int mtip_plan_slices(mtip_plan *p, const double *quats, int64_t N_o, const double *q,
                                int64_t N_p, const int32_t *pixel_order, int64_t n_order,
                                int tile_points, const int32_t *orient_order, double scale,
                                float *out, int64_t ld_out,
                                uint16_t *out_bf16, double *row_sum, double *row_sumsq,
                                int *bad_count, void *stream)
{
    MTIP_REQUIRE(p, MTIP_ERR_ARG, "mtip_plan_slices: null plan");
    MTIP_REQUIRE(p->loaded, MTIP_ERR_STATE, "mtip_plan_slices: no model set");
    MTIP_REQUIRE(N_o >= 0 && N_p >= 0, MTIP_ERR_ARG, "mtip_plan_slices: negative size");
    MTIP_REQUIRE(ld_out >= N_p, MTIP_ERR_ARG, "mtip_plan_slices: ld_out < N_p");
    MTIP_REQUIRE(N_o <= 0x7fffffff, MTIP_ERR_ARG, "mtip_plan_slices: N_o too large");
    MTIP_REQUIRE(std::isfinite(scale), MTIP_ERR_ARG, "mtip_plan_slices: scale not finite");
    if (N_o == 0 || N_p == 0) return MTIP_OK;
    MTIP_REQUIRE(quats && q && out, MTIP_ERR_ARG, "mtip_plan_slices: null pointer");
    MTIP_REQUIRE(N_p <= 0x7fffffff, MTIP_ERR_ARG, "mtip_plan_slices: N_p too large");
    if (tile_points == 0) tile_points = 256;
    MTIP_REQUIRE(tile_points == 64 || tile_points == 128 || tile_points == 256, MTIP_ERR_ARG,
                 "mtip_plan_slices: tile_points must be 0, 64, 128 or 256");
    MTIP_REQUIRE(!pixel_order || (n_order > 0 && n_order % tile_points == 0), MTIP_ERR_ARG,
                 "mtip_plan_slices: n_order must be a positive multiple of tile_points");
    DeviceGuard dg(p->device);
    GridView gv = view_of(p);
    // grid units per q unit: (pi / (E os)) * n / (2 pi)
    double gscale = scale * (double)p->n / (2.0 * M_PI);
    MTIP_W_DISPATCH(p->w, launch_slices, gv, quats, N_o, q, N_p, pixel_order, n_order,
                    tile_points, orient_order, gscale, out, ld_out, out_bf16, row_sum, row_sumsq,
                    bad_count, (hipStream_t)stream);
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}

extern "C" int mtip_rotate_positions(const double *quats, int64_t N_o, const double *q, int64_t N_p,
                                     double scale, double *H, double *K, double *L, void *stream)
{
    MTIP_REQUIRE(N_o >= 0 && N_p >= 0, MTIP_ERR_ARG, "mtip_rotate_positions: negative size");
    if (N_o == 0 || N_p == 0) return MTIP_OK;
    MTIP_REQUIRE(N_o <= 65535, MTIP_ERR_ARG, "mtip_rotate_positions: at most 65535 orientations per call");
    MTIP_REQUIRE(quats && q && H && K && L, MTIP_ERR_ARG, "mtip_rotate_positions: null pointer");
    hipLaunchKernelGGL(k_positions, dim3((unsigned)mtip::ceil_div(N_p, 256), (unsigned)N_o), dim3(256),
                       0, (hipStream_t)stream, quats, q, N_p, scale, H, K, L);
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}
