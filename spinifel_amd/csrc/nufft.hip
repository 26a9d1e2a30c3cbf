// Type-2 3D NUFFT of a real cubic model on gfx950: plan, model load (fine
// grid) and the fused rotation + interpolation slice kernel.
//
// Method (exponential-of-semicircle kernel, upsampling sigma = 2):
//   phi(z) = exp(beta (sqrt(1 - (2 z / w)^2) - 1)),  |z| < w / 2
//   b_l    = sum_k g_k exp(-2 pi i k . l / n),  g_k = sym(a)_k / prod phihat(k_d)
//   Re c_j = sum_l b_l prod_d phi(X_jd - l_d),  X = x n / (2 pi)
// sym(a)_k = (a(k) + a(-k)) / 2 makes b real (the reference keeps only the
// real part of the transform), so the fine grid is a real float32 array of
// (n + 2 pad)^3 values, indices l in [-n/2 - pad, n/2 + pad), evaluated
// directly at the padded indices so no wrap is needed while interpolating.
#include <math.h>
#include <vector>

#include "mtip_internal.h"

struct mtip_plan {
    int device = 0;
    int M = 0, h = 0, K = 0;      // model edge, half width, symmetric extent 2h+1
    int n = 0, w = 0, pad = 0, np = 0, off = 0;
    double beta = 0.0, eps = 0.0;
    float *grid = nullptr;        // (np)^3 real fine grid
    double *dinv = nullptr;       // K: 1 / phihat(k)
    double2 *tw = nullptr;        // n: exp(-2 pi i m / n)
    double2 *t1 = nullptr;        // K * K * np
    double2 *t2 = nullptr;        // K * np * np
    bool loaded = false;
};

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

constexpr int kOch = 8;    // outputs per thread along the transformed axis

// One separable DFT pass: out[a][o][c] = sum_k in(a, k, c) * tw((k * l(o)) mod n),
// k in [-h, h], l(o) = o - off.  MODE 0 reads the model (axis 2, C = 1) with
// symmetrisation + deconvolution; MODE 1 is complex -> complex; MODE 2
// writes the real part as float.
template <int MODE>
__global__ __launch_bounds__(256) void k_dft_pass(const float *__restrict__ ac,
                                                  const double2 *__restrict__ in,
                                                  double2 *__restrict__ out_c,
                                                  float *__restrict__ out_r,
                                                  const double *__restrict__ dinv,
                                                  const double2 *__restrict__ tw, int M, int h,
                                                  int n, int np, int off, int64_t A, int64_t C)
{
    extern __shared__ double2 s_tw[];
    for (int i = threadIdx.x; i < n; i += blockDim.x) s_tw[i] = tw[i];
    __syncthreads();

    const int K = 2 * h + 1;
    const int nchunk = (np + kOch - 1) / kOch;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A * nchunk * C) return;
    int64_t c = t % C;
    int64_t r = t / C;
    int chunk = (int)(r % nchunk);
    int64_t a = r / nchunk;
    int o0 = chunk * kOch;

    int lmod[kOch];
#pragma unroll
    for (int j = 0; j < kOch; ++j) {
        int l = o0 + j - off;
        lmod[j] = ((l % n) + n) % n;
    }
    double acc_re[kOch], acc_im[kOch];
#pragma unroll
    for (int j = 0; j < kOch; ++j) acc_re[j] = acc_im[j] = 0.0;

    for (int kk = 0; kk < K; ++kk) {
        int k = kk - h;
        double vr, vi;
        if (MODE == 0) {
            // a = k1 * K + k2 ; transformed axis is k3
            int k1 = (int)(a / K) - h, k2 = (int)(a % K) - h, k3 = k;
            int lo = -(M / 2), hi = (M - 1) / 2;
            double s = 0.0;
            if (k1 >= lo && k1 <= hi && k2 >= lo && k2 <= hi && k3 >= lo && k3 <= hi)
                s += ac[((int64_t)(k1 + M / 2) * M + (k2 + M / 2)) * M + (k3 + M / 2)];
            if (-k1 >= lo && -k1 <= hi && -k2 >= lo && -k2 <= hi && -k3 >= lo && -k3 <= hi)
                s += ac[((int64_t)(-k1 + M / 2) * M + (-k2 + M / 2)) * M + (-k3 + M / 2)];
            vr = 0.5 * s * dinv[k1 + h] * dinv[k2 + h] * dinv[k3 + h];
            vi = 0.0;
        } else {
            double2 v = in[(a * K + kk) * C + c];
            vr = v.x;
            vi = v.y;
        }
        int kmod = ((k % n) + n) % n;
#pragma unroll
        for (int j = 0; j < kOch; ++j) {
            double2 e = s_tw[(kmod * lmod[j]) % n];
            acc_re[j] += vr * e.x - vi * e.y;
            acc_im[j] += vr * e.y + vi * e.x;
        }
    }
#pragma unroll
    for (int j = 0; j < kOch; ++j) {
        int o = o0 + j;
        if (o < np) {
            int64_t idx = (a * np + o) * C + c;
            if (MODE == 2)
                out_r[idx] = (float)acc_re[j];
            else
                out_c[idx] = make_double2(acc_re[j], acc_im[j]);
        }
    }
}

// ------------------------------------------------------------ interpolation

struct GridView {
    const float *__restrict__ g;
    int np, off, n;
    float beta, c2;    // c2 = 4 / w^2
};

// phi(z) for the ES kernel; 0 outside the support.
__device__ __forceinline__ float es_phi(float z, float beta, float c2)
{
    float v = 1.0f - c2 * z * z;
    return v > 0.0f ? expf(beta * (sqrtf(v) - 1.0f)) : 0.0f;
}

// Interpolate Re c at grid coordinates X (grid units, already checked finite
// and |X| < 1.5 n).  Wraps into [-n/2, n/2), clamps the window start so every
// access stays inside the padded grid.
template <int W>
__device__ __forceinline__ float interp_point(const GridView &gv, double X0, double X1, double X2)
{
    const double dn = (double)gv.n;
    double X[3] = {X0, X1, X2};
    int i0[3];
    float ker[3][W];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double x = X[d] - dn * floor(X[d] / dn + 0.5);
        int s = (int)ceil(x - 0.5 * W);
        s = max(s, -gv.n / 2 - W);
        s = min(s, gv.n / 2 + 1);
        i0[d] = s + gv.off;
        float t = (float)(x - (double)s);
#pragma unroll
        for (int a = 0; a < W; ++a) ker[d][a] = es_phi((float)a - t, gv.beta, gv.c2);
    }
    const int64_t np = gv.np;
    float acc = 0.0f;
#pragma unroll
    for (int a = 0; a < W; ++a) {
        const float *plane = gv.g + (int64_t)(i0[0] + a) * np * np;
        float acc_b = 0.0f;
#pragma unroll
        for (int b = 0; b < W; ++b) {
            const float *row = plane + (int64_t)(i0[1] + b) * np + i0[2];
            float s = 0.0f;
#pragma unroll
            for (int c = 0; c < W; ++c) s = fmaf(ker[2][c], row[c], s);
            acc_b = fmaf(ker[1][b], s, acc_b);
        }
        acc = fmaf(ker[0][a], acc_b, acc);
    }
    return acc;
}

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

template <int W>
__global__ __launch_bounds__(256) void k_execute(GridView gv, const double *__restrict__ x,
                                                 const double *__restrict__ y,
                                                 const double *__restrict__ z, int64_t N,
                                                 float *__restrict__ out, int *bad)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const double g = (double)gv.n / (2.0 * M_PI);
    double X0 = x[j] * g, X1 = y[j] * g, X2 = z[j] * g;
    const double lim = 1.5 * gv.n;
    if (!(fabs(X0) < lim && fabs(X1) < lim && fabs(X2) < lim)) {
        out[j] = __builtin_nanf("");
        if (bad) atomicAdd(bad, 1);
        return;
    }
    out[j] = interp_point<W>(gv, X0, X1, X2);
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

// One workgroup per orientation: rotate every pixel q-vector in FP64, split
// into window start + FP32 fraction, interpolate, write FP32 (+ bf16) and the
// orientation's FP64 sum / sum of squares.
template <int W>
__global__ __launch_bounds__(256) void k_slices(GridView gv, const double *__restrict__ quats,
                                                const double *__restrict__ q, int64_t N_p,
                                                double gscale, float *__restrict__ out,
                                                int64_t ld_out, uint16_t *__restrict__ out_bf16,
                                                double *__restrict__ row_sum,
                                                double *__restrict__ row_sumsq, int *bad)
{
    __shared__ double s_red[8];
    const int64_t o = blockIdx.x;
    double ri[3][3];
    quat_inv_rot(quats + 4 * o, ri);
    const double lim = 1.5 * gv.n;
    double sum = 0.0, sumsq = 0.0;
    int nbad = 0;
    for (int64_t j = threadIdx.x; j < N_p; j += blockDim.x) {
        double qx = q[j], qy = q[N_p + j], qz = q[2 * N_p + j];
        double X0 = (ri[0][0] * qx + ri[0][1] * qy + ri[0][2] * qz) * gscale;
        double X1 = (ri[1][0] * qx + ri[1][1] * qy + ri[1][2] * qz) * gscale;
        double X2 = (ri[2][0] * qx + ri[2][1] * qy + ri[2][2] * qz) * gscale;
        float v;
        if (fabs(X0) < lim && fabs(X1) < lim && fabs(X2) < lim) {
            v = interp_point<W>(gv, X0, X1, X2);
        } else {
            v = __builtin_nanf("");
            ++nbad;
        }
        out[o * ld_out + j] = v;
        if (out_bf16) out_bf16[o * ld_out + j] = f32_to_bf16_rne(v);
        sum += (double)v;
        sumsq += (double)v * (double)v;
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
                   int64_t N_p, double gscale, float *out, int64_t ld, uint16_t *obf,
                   double *rs, double *rq, int *bad, hipStream_t st)
{
    hipLaunchKernelGGL(k_slices<W>, dim3((unsigned)N_o), dim3(256), 0, st, gv, quats, q, N_p,
                       gscale, out, ld, obf, rs, rq, bad);
}

template <int W>
void launch_execute(const GridView &gv, const double *x, const double *y, const double *z,
                    int64_t N, float *out, int *bad, hipStream_t st)
{
    hipLaunchKernelGGL(k_execute<W>, dim3((unsigned)mtip::ceil_div(N, 256)), dim3(256), 0, st,
                       gv, x, y, z, N, out, bad);
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
    gv.beta = (float)p->beta;
    gv.c2 = 4.0f / (float)(p->w * p->w);
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

extern "C" int mtip_plan_create(int M, double eps, int device, mtip_plan **out)
{
    MTIP_REQUIRE(out, MTIP_ERR_ARG, "mtip_plan_create: null output");
    *out = nullptr;
    MTIP_REQUIRE(M >= 2 && M <= 1024, MTIP_ERR_ARG, "mtip_plan_create: M must be in [2, 1024]");
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
    p->pad = w + 1;
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
    if (err == hipSuccess) err = hipMalloc(&p->t1, (size_t)p->K * p->K * p->np * sizeof(double2));
    if (err == hipSuccess) err = hipMalloc(&p->t2, (size_t)p->K * p->np * p->np * sizeof(double2));
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
    (void)hipFree(p->t1);
    (void)hipFree(p->t2);
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
    const int K = p->K, np = p->np, nch = (np + kOch - 1) / kOch;
    const size_t shm = p->n * sizeof(double2);
    // pass 0: axis 2 (k3 -> o3); A = K*K, C = 1
    int64_t A = (int64_t)K * K, C = 1;
    int64_t nt = A * nch * C;
    hipLaunchKernelGGL(k_dft_pass<0>, dim3((unsigned)mtip::ceil_div(nt, 256)), dim3(256), shm, st,
                       ac, nullptr, p->t1, nullptr, p->dinv, p->tw, p->M, p->h, p->n, np, p->off,
                       A, C);
    // pass 1: axis 1 (k2 -> o2); A = K, C = np
    A = K;
    C = np;
    nt = A * nch * C;
    hipLaunchKernelGGL(k_dft_pass<1>, dim3((unsigned)mtip::ceil_div(nt, 256)), dim3(256), shm, st,
                       nullptr, p->t1, p->t2, nullptr, p->dinv, p->tw, p->M, p->h, p->n, np,
                       p->off, A, C);
    // pass 2: axis 0 (k1 -> o1); A = 1, C = np * np ; real output
    A = 1;
    C = (int64_t)np * np;
    nt = A * nch * C;
    hipLaunchKernelGGL(k_dft_pass<2>, dim3((unsigned)mtip::ceil_div(nt, 256)), dim3(256), shm, st,
                       nullptr, p->t2, nullptr, p->grid, p->dinv, p->tw, p->M, p->h, p->n, np,
                       p->off, A, C);
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
                                int64_t N_p, double scale, float *out, int64_t ld_out,
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
    DeviceGuard dg(p->device);
    GridView gv = view_of(p);
    // grid units per q unit: (pi / (E os)) * n / (2 pi)
    double gscale = scale * (double)p->n / (2.0 * M_PI);
    MTIP_W_DISPATCH(p->w, launch_slices, gv, quats, N_o, q, N_p, gscale, out, ld_out, out_bf16,
                    row_sum, row_sumsq, bad_count, (hipStream_t)stream);
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
