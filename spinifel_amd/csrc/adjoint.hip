// Type-1 3D NUFFT (the adjoint of slicing) on gfx950: spreading onto the
// plan's fine grid and the separable transform back to the M^3 modes.
//
//   f_k = sum_j c_j exp(+i k . x_j),   k_d in [-M/2, (M-1)/2]
//
// Method (the plan's ES kernel phi, width w, fine grid n = next235even(2M)):
//   b_l   = sum_j c_j prod_d phi(l_d - X_jd),  X = x n / (2 pi)       (spread)
//   f~_k  = sum_l b_l exp(+2 pi i k . l / n)                           (passes)
//   f_k   = f~_k / prod_d phihat(k_d)
// Only Re f_k is produced: every caller keeps the real part (spinifel's
// adjoint with use_recip_sym, autocorrelation.py:123-124 and
// extern/nufft_ext.py:323-324).  For real c the grid b is real, so
// f~(-k) = conj f~(k); the passes transform the fast axis to k3 in [0, M/2]
// only and the last pass writes each (k1, k2, k3 > 0) result also at
// (-k1, -k2, -k3).
//
// b lives on the padded grid of the plan (indices l in [-n/2 - pad,
// n/2 + pad)), so windows never wrap; the passes sum over the padded indices,
// which folds the pads onto the periodic grid because exp(2 pi i k l / n)
// has period n in l.
//
// Spreading: a workgroup takes a tile of points (a detector tile of one
// orientation, or 128 consecutive caller points), bounds the union of their
// windows, accumulates the tile in an LDS box and adds the box's non-zero
// cells to the FP64 grid with FP64 global atomics (FP64 keeps the sum over
// all tiles exact to double rounding; each LDS cell sums at most a few
// hundred terms).  A tile whose box exceeds kBoxCap adds its taps to the
// FP64 grid directly.
#include <math.h>

#include "nufft_shared.h"

namespace {

constexpr int kAdjOch = 8;         // outputs per thread of a transform pass

__global__ __launch_bounds__(256) void k_zero_f64(double *__restrict__ p, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0.0;
}

// GEOM_ROTATE: workgroup o spreads row o of `vals` at the N_p pixel q-vectors
// rotated by inv(R(quats[o])) and scaled by gscale (grid units), tile after
// tile of TP points in the caller's pixel order.  GEOM_POINTS: workgroup b
// spreads points [TP b, TP b + TP) of (px, py, pz) (radians; gscale =
// n / 2 pi).  vals == nullptr spreads ones.
//
// Per tile: the TP window corners bound a box of the fine grid; when it fits
// kBoxCap the workgroup accumulates the tile there with plain LDS
// read-add-writes and then adds the box's non-zero cells to the FP64 grid.
// No LDS atomics (they run about one lane per clock on this part): wave v
// owns the box planes whose local x index is v mod 4, and each wave walks all
// of the tile's points in order, its lanes covering distinct (y, z) cells of
// the window (lane = row * W + element), so no two lanes ever update one cell
// at once and a wave's LDS accesses complete in issue order.  A box beyond
// kBoxCap sends the same per-lane taps to the FP64 grid as global atomics.
template <int W, int TP>
__global__ __launch_bounds__(256, 3) void k_spread(GridView gv, double *__restrict__ acc, int geom,
                                                 const double *__restrict__ quats,
                                                 const double *__restrict__ q, int64_t N_p,
                                                 const int32_t *__restrict__ order, int64_t n_order,
                                                 const double *__restrict__ px,
                                                 const double *__restrict__ py,
                                                 const double *__restrict__ pz, int64_t N_pts,
                                                 double gscale, const float *__restrict__ vals,
                                                 int64_t ld_v, int *bad)
{
    constexpr int RP = W * W <= 64 ? W : 64 / W;     // window rows per lane pass
    constexpr int NPASS = (W + RP - 1) / RP;
    // small tiles keep each point's 3 W kernel weights (the z ones times the
    // value) in LDS, computed once per tile instead of by every wave; their
    // room comes out of the box so three workgroups still fit a CU
    constexpr bool PRE = TP <= 128;
    constexpr int NW = 3 * W;
    constexpr int CAP = PRE ? kBoxCap - TP * NW : kBoxCap;
    __shared__ float s_box[CAP];
    __shared__ float s_w[PRE ? TP * NW : 1];
    __shared__ int s_base[TP];          // window corner: box-local or padded-grid index, -1 none
    __shared__ int s_xc[TP];            // box-local x of the corner
    __shared__ float4 s_pt[TP];         // fractional offsets tx, ty, tz and the value
    __shared__ int s_red_i[4][6];
    __shared__ double s_ri[9];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t o = blockIdx.x;
    const int np = gv.np;
    if (geom == GEOM_ROTATE) {
        if (tid == 0) {
            double ri[3][3];
            quat_inv_rot(quats + 4 * o, ri);
#pragma unroll
            for (int e = 0; e < 9; ++e) s_ri[e] = ri[e / 3][e % 3];
        }
        __syncthreads();
    }
    // this lane's window row and element (rows pass * RP + lrow)
    const int lrow = lane / W, lel = lane % W;
    const bool lane_on = lane < RP * W;
    int nbad = 0;
    const int64_t n_slots =
        geom == GEOM_POINTS ? TP : (order ? n_order : mtip::ceil_div(N_p, TP) * TP);

    for (int64_t tb = 0; tb < n_slots; tb += TP) {
        // ---- this thread's point (threads < TP)
        bool ok = false;
        float v = 0.f;
        int ijk[3] = {0, 0, 0};
        float t[3] = {0.f, 0.f, 0.f};
        if (tid < TP) {
            const int64_t k = tb + tid;
            double X0 = 0.0, X1 = 0.0, X2 = 0.0;
            bool have = false;
            if (geom == GEOM_POINTS) {
                const int64_t j = o * TP + tid;
                if (j < N_pts) {
                    have = true;
                    X0 = px[j] * gscale;
                    X1 = py[j] * gscale;
                    X2 = pz[j] * gscale;
                    v = vals ? vals[j] : 1.f;
                }
            } else {
                const int64_t pix = order ? (k < n_order ? (int64_t)order[k] : -1)
                                          : (k < N_p ? k : -1);
                if (pix >= 0) {
                    have = true;
                    const double x = q[pix], y = q[N_p + pix], z = q[2 * N_p + pix];
                    X0 = (s_ri[0] * x + s_ri[1] * y + s_ri[2] * z) * gscale;
                    X1 = (s_ri[3] * x + s_ri[4] * y + s_ri[5] * z) * gscale;
                    X2 = (s_ri[6] * x + s_ri[7] * y + s_ri[8] * z) * gscale;
                    v = vals ? vals[o * ld_v + pix] : 1.f;
                }
            }
            if (have) {
                ok = point_corner<W>(gv, X0, X1, X2, ijk, t);
                if (!ok) ++nbad;
                if (v == 0.f) ok = false;     // contributes nothing
            }
        }
        // ---- box of the tile's windows (block min / max)
        {
            int r[6] = {ok ? ijk[0] : 0x7fffffff, ok ? ijk[1] : 0x7fffffff, ok ? ijk[2] : 0x7fffffff,
                        ok ? -ijk[0] : 0x7fffffff, ok ? -ijk[1] : 0x7fffffff, ok ? -ijk[2] : 0x7fffffff};
#pragma unroll
            for (int e = 0; e < 6; ++e) r[e] = wave_min_i(r[e]);
            if (lane == 0)
#pragma unroll
                for (int e = 0; e < 6; ++e) s_red_i[wave][e] = r[e];
        }
        __syncthreads();
        int lo[3], hi[3];
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            lo[e] = min(min(s_red_i[0][e], s_red_i[1][e]), min(s_red_i[2][e], s_red_i[3][e]));
            hi[e] = -min(min(s_red_i[0][e + 3], s_red_i[1][e + 3]),
                         min(s_red_i[2][e + 3], s_red_i[3][e + 3]));
        }
        if (lo[0] == 0x7fffffff) {
            __syncthreads();     // s_red_i is rewritten by the next tile
            continue;
        }
        const int bx = hi[0] - lo[0] + W, by = hi[1] - lo[1] + W, bz = hi[2] - lo[2] + W;
        const bool staged = bx * by * bz <= CAP;
        const int ncell = staged ? bx * by * bz : 0;
        if (tid < TP) {
            s_base[tid] = !ok ? -1
                        : staged ? ((ijk[0] - lo[0]) * by + (ijk[1] - lo[1])) * bz + (ijk[2] - lo[2])
                                 : (ijk[0] * np + ijk[1]) * np + ijk[2];
            s_xc[tid] = ijk[0] - lo[0];
            s_pt[tid] = make_float4(t[0], t[1], t[2], v);
        }
        for (int k = tid; k < ncell; k += 256) s_box[k] = 0.f;
        __syncthreads();
        if constexpr (PRE) {
            for (int k = tid; k < TP * NW; k += 256) {
                const int pn = k / NW, r = k - pn * NW, d = r / W, e = r - d * W;
                const float4 pt = s_pt[pn];
                const float tt = d == 0 ? pt.x : d == 1 ? pt.y : pt.z;
                const float w = es_phi((float)e - tt, gv.beta_l2e, gv.c2);
                s_w[k] = d == 2 ? w * pt.w : w;
            }
            __syncthreads();
        }
        // ---- taps: every wave walks all TP points, updating its own x planes;
        // a chunk of 64 points' parameters is read into registers at once and
        // each point's taken from its lane by readlane (scalar operands).  A
        // point's planes (a0, a0 + 4, ...) are read together, then added and
        // written, so one LDS round trip serves them all; with precomputed
        // weights the next point's are read while this point's planes wait.
        const int plane = staged ? by * bz : np * np, rowstride = staged ? bz : np;
        constexpr int NPL = (W + 3) / 4;      // planes per point and wave, at most
        auto taps = [&](int base, int a0, float pa, const float (&wyz)[NPASS])
            __attribute__((always_inline)) {
#pragma unroll
            for (int ps = 0; ps < NPASS; ++ps) {
                const int b = ps * RP + lrow;
                if (!lane_on || b >= W) continue;
                const int off = base + b * rowstride + lel + a0 * plane;
                float wa[NPL];
#pragma unroll
                for (int j = 0; j < NPL; ++j)
                    wa[j] = a0 + 4 * j < W ? __int_as_float(__builtin_amdgcn_readlane(
                                                 __float_as_int(pa), a0 + 4 * j))
                                           : 0.f;
                if (staged) {
                    float cur[NPL];
#pragma unroll
                    for (int j = 0; j < NPL; ++j)
                        if (a0 + 4 * j < W) cur[j] = s_box[off + 4 * j * plane];
#pragma unroll
                    for (int j = 0; j < NPL; ++j)
                        if (a0 + 4 * j < W) s_box[off + 4 * j * plane] = fmaf(wa[j], wyz[ps], cur[j]);
                } else {
#pragma unroll
                    for (int j = 0; j < NPL; ++j)
                        if (a0 + 4 * j < W)
                            unsafeAtomicAdd(acc + (int64_t)off + (int64_t)(4 * j) * plane,
                                            (double)(wa[j] * wyz[ps]));
                }
            }
        };
#pragma unroll 1
        for (int c0 = 0; c0 < TP; c0 += 64) {
            const int cl = c0 + lane < TP ? c0 + lane : TP - 1;
            const int r_base = s_base[cl], r_xc = s_xc[cl];
            const int nc = TP - c0 < 64 ? TP - c0 : 64;
            if constexpr (PRE) {
                // weights of point i: x by lane (lanes < W), y by row, z (x value) by element
                auto load = [&](int i, float &pa, float (&wy)[NPASS], float &wz)
                    __attribute__((always_inline)) {
                    const float *wp = s_w + (c0 + i) * NW;
                    pa = wp[lane < W ? lane : 0];
                    wz = wp[2 * W + lel];
#pragma unroll
                    for (int ps = 0; ps < NPASS; ++ps) {
                        const int b = ps * RP + lrow;
                        wy[ps] = wp[W + (b < W ? b : W - 1)];
                    }
                };
                float n_pa, n_wy[NPASS], n_wz;
                load(0, n_pa, n_wy, n_wz);
                for (int i = 0; i < nc; ++i) {
                    const float pa = n_pa, wz = n_wz;
                    float wyz[NPASS];
#pragma unroll
                    for (int ps = 0; ps < NPASS; ++ps) wyz[ps] = n_wy[ps] * wz;
                    if (i + 1 < nc) load(i + 1, n_pa, n_wy, n_wz);
                    const int base = __builtin_amdgcn_readlane(r_base, i);
                    if (base < 0) continue;
                    taps(base, (wave - __builtin_amdgcn_readlane(r_xc, i)) & 3, pa, wyz);
                }
            } else {
                const float4 r_pt = s_pt[cl];
                for (int i = 0; i < nc; ++i) {
                    const int base = __builtin_amdgcn_readlane(r_base, i);
                    if (base < 0) continue;
                    const float tx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r_pt.x), i));
                    const float ty = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r_pt.y), i));
                    const float tz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r_pt.z), i));
                    const float pv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r_pt.w), i));
                    // lanes 0 .. W-1 hold the x weights, read back per plane
                    const float pa = es_phi((float)lane - tx, gv.beta_l2e, gv.c2);
                    const float wz = es_phi((float)lel - tz, gv.beta_l2e, gv.c2) * pv;
                    float wyz[NPASS];
#pragma unroll
                    for (int ps = 0; ps < NPASS; ++ps)
                        wyz[ps] = es_phi((float)(ps * RP + lrow) - ty, gv.beta_l2e, gv.c2) * wz;
                    taps(base, (wave - __builtin_amdgcn_readlane(r_xc, i)) & 3, pa, wyz);
                }
            }
        }
        __syncthreads();
        // ---- the box's non-zero cells into the FP64 grid
        if (staged) {
            const float inv_bz = 1.0f / (float)bz, inv_by = 1.0f / (float)by;
            for (int k = tid; k < ncell; k += 256) {
                const float val = s_box[k];
                if (val != 0.f) {
                    // k / bz and r / by for k < 2^16: the float estimates are
                    // within half a unit of the integer quotient
                    const int r = (int)(((float)k + 0.5f) * inv_bz), z = k - r * bz;
                    const int xr = (int)(((float)r + 0.5f) * inv_by), yr = r - xr * by;
                    unsafeAtomicAdd(acc + ((int64_t)(lo[0] + xr) * np + (lo[1] + yr)) * np + lo[2] + z,
                                    (double)val);
                }
            }
        }
        __syncthreads();
    }
    if (nbad && bad) atomicAdd(bad, nbad);
}

// One separable pass of the transform back to modes:
//   out[a][kk][c] = sum_{o < L} in(a, o, c) exp(+2 pi i k l / n),
//   k = kk + k_lo (kk < NK), l = o - off.
// MODE 0: real input, complex output.  MODE 1: complex -> complex.
// MODE 2 (last pass, axis 0, A = 1, c = k2i (h + 1) + k3): writes
//   out_r[i1][i2][i3] = Re(.) dinv(k1) dinv(k2) dinv(k3) scale support[i]
// at (k1, k2, k3) and, for k3 > 0, at (-k1, -k2, -k3), where i_d = k_d + M/2.
template <int MODE>
__global__ __launch_bounds__(256) void k_adj_pass(const double *__restrict__ in_r,
                                                  const double2 *__restrict__ in_c,
                                                  double2 *__restrict__ out_c,
                                                  const double2 *__restrict__ tw, int n, int L,
                                                  int off, int NK, int k_lo, int64_t A, int64_t C,
                                                  double *__restrict__ out_r,
                                                  const double *__restrict__ dinv, int h, int M,
                                                  const double *__restrict__ support, double scale)
{
    extern __shared__ double2 s_tw[];
    for (int i = threadIdx.x; i < n; i += blockDim.x) s_tw[i] = tw[i];
    __syncthreads();

    const int nchunk = (NK + kAdjOch - 1) / kAdjOch;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A * nchunk * C) return;
    const int64_t c = t % C;
    const int64_t r = t / C;
    const int chunk = (int)(r % nchunk);
    const int64_t a = r / nchunk;
    const int kk0 = chunk * kAdjOch;

    // phase index (k l) mod n of each output, advanced by k mod n per input
    const int l0 = ((-off) % n + n) % n;
    int idx[kAdjOch], step[kAdjOch];
#pragma unroll
    for (int j = 0; j < kAdjOch; ++j) {
        const int k = kk0 + j + k_lo;
        step[j] = ((k % n) + n) % n;
        idx[j] = (int)(((int64_t)step[j] * l0) % n);
    }
    double acc_re[kAdjOch], acc_im[kAdjOch];
#pragma unroll
    for (int j = 0; j < kAdjOch; ++j) acc_re[j] = acc_im[j] = 0.0;

    for (int o = 0; o < L; ++o) {
        double vr, vi = 0.0;
        if (MODE == 0) {
            vr = in_r[a * L + o];
        } else {
            const double2 v = in_c[(a * L + o) * C + c];
            vr = v.x;
            vi = v.y;
        }
#pragma unroll
        for (int j = 0; j < kAdjOch; ++j) {
            // tw = exp(-i theta): exp(+i theta) = (tw.x, -tw.y)
            const double2 e = s_tw[idx[j]];
            const double cs = e.x, sn = -e.y;
            if (MODE == 0) {
                acc_re[j] = fma(vr, cs, acc_re[j]);
                acc_im[j] = fma(vr, sn, acc_im[j]);
            } else {
                acc_re[j] = fma(vr, cs, fma(-vi, sn, acc_re[j]));
                acc_im[j] = fma(vr, sn, fma(vi, cs, acc_im[j]));
            }
            idx[j] += step[j];
            idx[j] -= idx[j] >= n ? n : 0;
        }
    }
#pragma unroll
    for (int j = 0; j < kAdjOch; ++j) {
        const int kk = kk0 + j;
        if (kk >= NK) break;
        if (MODE != 2) {
            out_c[(a * NK + kk) * C + c] = make_double2(acc_re[j], acc_im[j]);
        } else {
            const int k1 = kk + k_lo;
            const int hp1 = h + 1;
            const int k2 = (int)(c / hp1) - h, k3 = (int)(c % hp1);
            const double val =
                acc_re[j] * dinv[k1 + h] * dinv[k2 + h] * dinv[k3 + h] * scale;
            const int lo = -(M / 2), hi = (M - 1) / 2;
            if (k1 <= hi && k2 <= hi && k3 <= hi) {
                const int64_t i = ((int64_t)(k1 - lo) * M + (k2 - lo)) * M + (k3 - lo);
                out_r[i] = support ? val * support[i] : val;
            }
            if (k3 > 0 && -k1 >= lo && -k1 <= hi && -k2 >= lo && -k2 <= hi && -k3 >= lo) {
                const int64_t i = ((int64_t)(-k1 - lo) * M + (-k2 - lo)) * M + (-k3 - lo);
                out_r[i] = support ? val * support[i] : val;
            }
        }
    }
}

template <int W>
void launch_spread(const GridView &gv, double *acc, int geom, const double *quats, int64_t N_o,
                   const double *q, int64_t N_p, const int32_t *order, int64_t n_order, int tile,
                   const double *x, const double *y, const double *z, int64_t N, double gscale,
                   const float *vals, int64_t ld, int *bad, hipStream_t st)
{
#define MTIP_SPREAD_TP(T)                                                                      \
    hipLaunchKernelGGL((k_spread<W, T>),                                                       \
                       dim3((unsigned)(geom == GEOM_ROTATE ? N_o : mtip::ceil_div(N, T))),     \
                       dim3(256), 0, st, gv, acc, geom, quats, q, N_p, order, n_order, x, y, z, \
                       N, gscale, vals, ld, bad)
    if (tile == 64) MTIP_SPREAD_TP(64);
    else if (tile == 128) MTIP_SPREAD_TP(128);
    else MTIP_SPREAD_TP(256);
#undef MTIP_SPREAD_TP
}

}  // namespace

extern "C" int mtip_plan_spread_begin(mtip_plan *p, void *stream)
{
    MTIP_REQUIRE(p, MTIP_ERR_ARG, "mtip_plan_spread_begin: null plan");
    DeviceGuard dg(p->device);
    MTIP_REQUIRE(dg.ok, MTIP_ERR_HIP, "mtip_plan_spread_begin: cannot select device");
    const int64_t np = p->np;
    if (!p->acc) {
        const size_t n_acc = (size_t)np * np * np;
        const size_t n_a1 = (size_t)np * np * (p->h + 1);
        const size_t n_a2 = (size_t)np * p->K * (p->h + 1);
        hipError_t err = hipMalloc(&p->acc, n_acc * sizeof(double));
        if (err == hipSuccess) err = hipMalloc(&p->a1, n_a1 * sizeof(double2));
        if (err == hipSuccess) err = hipMalloc(&p->a2, n_a2 * sizeof(double2));
        if (err != hipSuccess) {
            (void)hipFree(p->acc);
            (void)hipFree(p->a1);
            (void)hipFree(p->a2);
            p->acc = nullptr;
            p->a1 = p->a2 = nullptr;
            return mtip::fail(err == hipErrorOutOfMemory ? MTIP_ERR_NOMEM : MTIP_ERR_HIP,
                              std::string("mtip_plan_spread_begin: ") + hipGetErrorString(err));
        }
    }
    const int64_t n_acc = np * np * np;
    hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)std::min<int64_t>(mtip::ceil_div(n_acc, 256), 8192)),
                       dim3(256), 0, (hipStream_t)stream, p->acc, n_acc);
    MTIP_HIP_CHECK(hipGetLastError());
    p->spread_open = true;
    return MTIP_OK;
}

extern "C" int mtip_plan_spread_slices(mtip_plan *p, const double *quats, int64_t N_o,
                                       const double *q, int64_t N_p, const int32_t *pixel_order,
                                       int64_t n_order, int tile_points, double scale,
                                       const float *values,
                                       int64_t ld_values, int *bad_count, void *stream)
{
    MTIP_REQUIRE(p, MTIP_ERR_ARG, "mtip_plan_spread_slices: null plan");
    MTIP_REQUIRE(p->spread_open, MTIP_ERR_STATE,
                 "mtip_plan_spread_slices: mtip_plan_spread_begin not called");
    MTIP_REQUIRE(N_o >= 0 && N_p >= 0, MTIP_ERR_ARG, "mtip_plan_spread_slices: negative size");
    MTIP_REQUIRE(N_o <= 0x7fffffff && N_p <= 0x7fffffff, MTIP_ERR_ARG,
                 "mtip_plan_spread_slices: size too large");
    MTIP_REQUIRE(std::isfinite(scale), MTIP_ERR_ARG, "mtip_plan_spread_slices: scale not finite");
    if (N_o == 0 || N_p == 0) return MTIP_OK;
    MTIP_REQUIRE(quats && q, MTIP_ERR_ARG, "mtip_plan_spread_slices: null pointer");
    MTIP_REQUIRE(!values || ld_values >= N_p, MTIP_ERR_ARG,
                 "mtip_plan_spread_slices: ld_values < N_p");
    if (tile_points == 0) tile_points = 256;
    MTIP_REQUIRE(tile_points == 64 || tile_points == 128 || tile_points == 256, MTIP_ERR_ARG,
                 "mtip_plan_spread_slices: tile_points must be 0, 64, 128 or 256");
    MTIP_REQUIRE(!pixel_order || (n_order > 0 && n_order % tile_points == 0), MTIP_ERR_ARG,
                 "mtip_plan_spread_slices: n_order must be a positive multiple of tile_points");
    DeviceGuard dg(p->device);
    GridView gv = view_of(p);
    const double gscale = scale * (double)p->n / (2.0 * M_PI);
    MTIP_W_DISPATCH(p->w, launch_spread, gv, p->acc, (int)GEOM_ROTATE, quats, N_o, q, N_p,
                    pixel_order, n_order, tile_points, nullptr, nullptr, nullptr, (int64_t)0,
                    gscale, values,
                    ld_values, bad_count, (hipStream_t)stream);
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}

extern "C" int mtip_plan_spread_points(mtip_plan *p, const double *x, const double *y,
                                       const double *z, int64_t N, const float *values,
                                       int *bad_count, void *stream)
{
    MTIP_REQUIRE(p, MTIP_ERR_ARG, "mtip_plan_spread_points: null plan");
    MTIP_REQUIRE(p->spread_open, MTIP_ERR_STATE,
                 "mtip_plan_spread_points: mtip_plan_spread_begin not called");
    MTIP_REQUIRE(N >= 0, MTIP_ERR_ARG, "mtip_plan_spread_points: N < 0");
    if (N == 0) return MTIP_OK;
    MTIP_REQUIRE(x && y && z, MTIP_ERR_ARG, "mtip_plan_spread_points: null pointer");
    MTIP_REQUIRE(mtip::ceil_div(N, 128) <= 0x7fffffff, MTIP_ERR_ARG,
                 "mtip_plan_spread_points: N too large");
    DeviceGuard dg(p->device);
    GridView gv = view_of(p);
    const double gscale = (double)p->n / (2.0 * M_PI);
    // 128-point tiles: the kernel weights of a tile's points are computed once
    // into LDS (256-point tiles recompute them in every wave)
    MTIP_W_DISPATCH(p->w, launch_spread, gv, p->acc, (int)GEOM_POINTS, nullptr, (int64_t)0,
                    nullptr, (int64_t)0, nullptr, (int64_t)0, 128, x, y, z, N, gscale, values,
                    (int64_t)0, bad_count, (hipStream_t)stream);
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}

extern "C" int mtip_plan_spread_finish(mtip_plan *p, const double *support, double scale,
                                       double *out, void *stream)
{
    MTIP_REQUIRE(p && out, MTIP_ERR_ARG, "mtip_plan_spread_finish: null argument");
    MTIP_REQUIRE(p->spread_open, MTIP_ERR_STATE,
                 "mtip_plan_spread_finish: mtip_plan_spread_begin not called");
    MTIP_REQUIRE(std::isfinite(scale), MTIP_ERR_ARG, "mtip_plan_spread_finish: scale not finite");
    DeviceGuard dg(p->device);
    hipStream_t st = (hipStream_t)stream;
    const int np = p->np, h = p->h, K = p->K, n = p->n;
    const int nh = h + 1;
    const size_t shm = n * sizeof(double2);
    auto nblk = [](int64_t A, int NK, int64_t C) {
        return (unsigned)mtip::ceil_div(A * ((NK + kAdjOch - 1) / kAdjOch) * C, 256);
    };
    // pass 0: axis 2 (l3 -> k3 in [0, h]), real input [l1 l2][l3]
    int64_t A = (int64_t)np * np, C = 1;
    hipLaunchKernelGGL(k_adj_pass<0>, dim3(nblk(A, nh, C)), dim3(256), shm, st, p->acc, nullptr,
                       p->a1, p->tw, n, np, p->off, nh, 0, A, C, nullptr, nullptr, h, p->M,
                       nullptr, 0.0);
    // pass 1: axis 1 (l2 -> k2 in [-h, h]), input [l1][l2][k3]
    A = np;
    C = nh;
    hipLaunchKernelGGL(k_adj_pass<1>, dim3(nblk(A, K, C)), dim3(256), shm, st, nullptr, p->a1,
                       p->a2, p->tw, n, np, p->off, K, -h, A, C, nullptr, nullptr, h, p->M,
                       nullptr, 0.0);
    // pass 2: axis 0 (l1 -> k1 in [-h, h]), input [l1][k2 k3], real output
    A = 1;
    C = (int64_t)K * nh;
    hipLaunchKernelGGL(k_adj_pass<2>, dim3(nblk(A, K, C)), dim3(256), shm, st, nullptr, p->a2,
                       nullptr, p->tw, n, np, p->off, K, -h, A, C, out, p->dinv, h, p->M, support,
                       scale);
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}
