// The merge step's per-iteration vector work on gfx950 (FP64): the operand
// preparation and recombination around the FFTs of
//   W u = support . crop(irfft(rfft(pad(support . u)) . F)) + rlambda u
//         + flambda support . irfft(rfft(support . u) . mask)
// (spinifel/sequential/autocorrelation.py:110-145, :173-178), and the
// conjugate-gradient updates (cupyx/scipy cg, :211-212) with the step
// lengths kept on the device.  Dot products reduce in two fixed-order passes
// (partials per workgroup, then one workgroup), so results do not depend on
// scheduling.
#include "mtip_internal.h"

namespace {

constexpr int kDotBlocks = 512;

// ups[:M, :M, :M] = support . u (the rest of ups stays as the caller zeroed
// it once); us = support . u as an M^3 array for the regulariser's transform.
__global__ __launch_bounds__(256) void k_merge_pad(const double *__restrict__ u,
                                                   const double *__restrict__ s, int M, int Mu,
                                                   double *__restrict__ ups, double *__restrict__ us)
{
    const int64_t n = (int64_t)M * M * M;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = (int)(i % M), j = (int)((i / M) % M), l = (int)(i / ((int64_t)M * M));
    const double v = s ? u[i] * s[i] : u[i];
    ups[((int64_t)l * Mu + j) * Mu + k] = v;
    us[i] = v;
}

// out = s . y_ups[:M, :M, :M] + rlambda u + flambda s . y_reg
__global__ __launch_bounds__(256) void k_merge_combine(const double *__restrict__ y_ups,
                                                       const double *__restrict__ y_reg,
                                                       const double *__restrict__ u,
                                                       const double *__restrict__ s, int M,
                                                       int Mu, double rlambda, double flambda,
                                                       double *__restrict__ out)
{
    const int64_t n = (int64_t)M * M * M;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = (int)(i % M), j = (int)((i / M) % M), l = (int)(i / ((int64_t)M * M));
    const double sv = s ? s[i] : 1.0;
    const double ada = y_ups[((int64_t)l * Mu + j) * Mu + k] * sv;
    const double fdf = y_reg[i] * sv;
    out[i] = (ada + rlambda * u[i]) + flambda * fdf;
}

__device__ __forceinline__ double block_sum(double v, double *red)
{
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
    return t;
}

// partial[b] = sum over this workgroup's grid-stride elements of a . b
__global__ __launch_bounds__(256) void k_dot_partial(const double *__restrict__ a,
                                                     const double *__restrict__ b, int64_t n,
                                                     double *__restrict__ partial)
{
    __shared__ double red[4];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        acc = fma(a[i], b[i], acc);
    const double t = block_sum(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void k_dot_final(const double *__restrict__ partial, int n,
                                                   double *__restrict__ out)
{
    __shared__ double red[4];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
    const double t = block_sum(acc, red);
    if (threadIdx.x == 0) *out = t;
}

// x += (rho / pq) p;  r -= (rho / pq) q
__global__ __launch_bounds__(256) void k_cg_update(double *__restrict__ x, double *__restrict__ r,
                                                   const double *__restrict__ p,
                                                   const double *__restrict__ q,
                                                   const double *__restrict__ rho,
                                                   const double *__restrict__ pq, int64_t n)
{
    const double alpha = *rho / *pq;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        x[i] += alpha * p[i];
        r[i] -= alpha * q[i];
    }
}

// p = r (first) or p = r + (rho / rho_old) p
__global__ __launch_bounds__(256) void k_cg_direction(double *__restrict__ p,
                                                      const double *__restrict__ r,
                                                      const double *__restrict__ rho,
                                                      const double *__restrict__ rho_old,
                                                      int64_t n, int first)
{
    const double beta = first ? 0.0 : *rho / *rho_old;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = first ? r[i] : r[i] + beta * p[i];
}

unsigned blocks_for(int64_t n, int64_t cap)
{
    const int64_t b = mtip::ceil_div(n, 256);
    return (unsigned)(b < 1 ? 1 : (b > cap ? cap : b));
}

}  // namespace

extern "C" int mtip_merge_pad(const double *u, const double *support, int M, int M_ups,
                              double *ups, double *us, void *stream)
{
    MTIP_REQUIRE(M >= 1 && M_ups >= M, MTIP_ERR_ARG, "mtip_merge_pad: need 1 <= M <= M_ups");
    MTIP_REQUIRE(u && ups && us, MTIP_ERR_ARG, "mtip_merge_pad: null pointer");
    const int64_t n = (int64_t)M * M * M;
    hipLaunchKernelGGL(k_merge_pad, dim3((unsigned)mtip::ceil_div(n, 256)), dim3(256), 0,
                       (hipStream_t)stream, u, support, M, M_ups, ups, us);
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}

extern "C" int mtip_merge_combine(const double *y_ups, const double *y_reg, const double *u,
                                  const double *support, int M, int M_ups, double rlambda,
                                  double flambda, double *out, void *stream)
{
    MTIP_REQUIRE(M >= 1 && M_ups >= M, MTIP_ERR_ARG, "mtip_merge_combine: need 1 <= M <= M_ups");
    MTIP_REQUIRE(y_ups && y_reg && u && out, MTIP_ERR_ARG, "mtip_merge_combine: null pointer");
    const int64_t n = (int64_t)M * M * M;
    hipLaunchKernelGGL(k_merge_combine, dim3((unsigned)mtip::ceil_div(n, 256)), dim3(256), 0,
                       (hipStream_t)stream, y_ups, y_reg, u, support, M, M_ups, rlambda, flambda,
                       out);
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}

extern "C" size_t mtip_dot_workspace_size(void) { return kDotBlocks * sizeof(double); }

extern "C" int mtip_dot(const double *a, const double *b, int64_t n, double *work,
                        size_t work_bytes, double *out, void *stream)
{
    MTIP_REQUIRE(n >= 0, MTIP_ERR_ARG, "mtip_dot: n < 0");
    MTIP_REQUIRE(a && b && work && out, MTIP_ERR_ARG, "mtip_dot: null pointer");
    MTIP_REQUIRE(work_bytes >= mtip_dot_workspace_size(), MTIP_ERR_NOMEM,
                 "mtip_dot: workspace smaller than mtip_dot_workspace_size()");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_dot_partial, dim3(kDotBlocks), dim3(256), 0, st, a, b, n, work);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, st, work, kDotBlocks, out);
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}

extern "C" int mtip_cg_update(double *x, double *r, const double *p, const double *q,
                              const double *rho, const double *pq, int64_t n, void *stream)
{
    MTIP_REQUIRE(n >= 0, MTIP_ERR_ARG, "mtip_cg_update: n < 0");
    if (n == 0) return MTIP_OK;
    MTIP_REQUIRE(x && r && p && q && rho && pq, MTIP_ERR_ARG, "mtip_cg_update: null pointer");
    hipLaunchKernelGGL(k_cg_update, dim3(blocks_for(n, 4096)), dim3(256), 0, (hipStream_t)stream,
                       x, r, p, q, rho, pq, n);
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}

extern "C" int mtip_cg_direction(double *p, const double *r, const double *rho,
                                 const double *rho_old, int64_t n, int first, void *stream)
{
    MTIP_REQUIRE(n >= 0, MTIP_ERR_ARG, "mtip_cg_direction: n < 0");
    if (n == 0) return MTIP_OK;
    MTIP_REQUIRE(p && r && rho && (first || rho_old), MTIP_ERR_ARG,
                 "mtip_cg_direction: null pointer");
    hipLaunchKernelGGL(k_cg_direction, dim3(blocks_for(n, 4096)), dim3(256), 0,
                       (hipStream_t)stream, p, r, rho, rho_old, n, first);
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}
