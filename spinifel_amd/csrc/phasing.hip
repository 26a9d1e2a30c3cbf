// Phase-retrieval projections of spinifel's phasing step on gfx950 (FP64):
// the Fourier-amplitude replacement of step_phase on the half spectrum of a
// real density, and the fused ER / HIO real-space update.
//
// The density is real, so its spectrum is Hermitian; the amplitudes (of a
// real autocorrelation) are even and the phases odd, so the replaced
// spectrum stays Hermitian and the caller's real-to-complex transforms give
// exactly the real part the reference keeps
// (spinifel/sequential/phasing.py:189-192).
#include "mtip_internal.h"

namespace {

// X[i] <- amp[i] exp(i angle(X[i])) for every i but the zero frequency
// (amp_mask_[0, 0, 0] = 0, phasing.py:269-270); angle(0) = 0 as numpy's.
__global__ __launch_bounds__(256) void k_amp_project(double2 *__restrict__ X,
                                                     const double *__restrict__ amp, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        if (i == 0) continue;
        const double2 x = X[i];
        const double r = hypot(x.x, x.y);
        const double a = amp[i];
        X[i] = r > 0.0 ? make_double2(a * (x.x / r), a * (x.y / r)) : make_double2(a, 0.0);
    }
}

// ER (mode 0, phasing.py:141-146) and HIO (mode 1, :162-167) from the
// projected density rho_mod: star = support && rho_mod > 0, then
//   ER:  rho = star ? rho_mod : 0;             rho = rho_max where rho_mod > rho_max
//   HIO: rho = star ? rho_mod : rho - b rho_mod; rho += 2 b rho_mod - rho_max there
__global__ __launch_bounds__(256) void k_er_hio(double *__restrict__ rho,
                                                const double *__restrict__ rho_mod,
                                                const uint8_t *__restrict__ support, int64_t n,
                                                int mode, double beta, double rho_max)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        const double m = rho_mod[i];
        const bool star = support[i] != 0 && m > 0.0;
        double r;
        if (mode == 0) {
            r = star ? m : 0.0;
            if (m > rho_max) r = rho_max;
        } else {
            r = star ? m : rho[i] - beta * m;
            if (m > rho_max) r += 2.0 * beta * m - rho_max;
        }
        rho[i] = r;
    }
}

unsigned grid_for(int64_t n)
{
    int64_t b = mtip::ceil_div(n, 256);
    return (unsigned)(b < 16384 ? (b > 0 ? b : 1) : 16384);
}

}  // namespace

extern "C" int mtip_phase_amplitudes(void *spectrum, const double *amplitudes, int64_t n,
                                     void *stream)
{
    MTIP_REQUIRE(n >= 0, MTIP_ERR_ARG, "mtip_phase_amplitudes: n < 0");
    if (n == 0) return MTIP_OK;
    MTIP_REQUIRE(spectrum && amplitudes, MTIP_ERR_ARG, "mtip_phase_amplitudes: null pointer");
    hipLaunchKernelGGL(k_amp_project, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream,
                       (double2 *)spectrum, amplitudes, n);
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}

extern "C" int mtip_phase_update(double *rho, const double *rho_mod, const uint8_t *support,
                                 int64_t n, int mode, double beta, double rho_max, void *stream)
{
    MTIP_REQUIRE(n >= 0, MTIP_ERR_ARG, "mtip_phase_update: n < 0");
    MTIP_REQUIRE(mode == 0 || mode == 1, MTIP_ERR_ARG, "mtip_phase_update: mode must be 0 (ER) or 1 (HIO)");
    if (n == 0) return MTIP_OK;
    MTIP_REQUIRE(rho && rho_mod && support, MTIP_ERR_ARG, "mtip_phase_update: null pointer");
    hipLaunchKernelGGL(k_er_hio, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, rho,
                       rho_mod, support, n, mode, beta, rho_max);
    MTIP_HIP_CHECK(hipGetLastError());
    return MTIP_OK;
}
