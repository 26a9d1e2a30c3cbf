/*
 * ORACLE -- test infrastructure, not product code.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker / the timed CPU baseline.
 *
 * CPU restatement (FP64, OpenMP) of the interpolation step of the
 * exponential-of-semicircle (ES) type-2 3D NUFFT that spinifel calls through
 * finufftpy 1.1.2 `nufft3d2(x, y, z, c, isign, eps, f)`
 * (spinifel/extern/finufft_ext.py:70-94, call at :88; version gate
 * spinifel/extern/__init__.py:79-87).  finufft itself is not vendored in the
 * reference (empty submodule, .gitmodules:10-12); this follows its published
 * method (Barnett, Magland, af Klinteberg, SIAM J. Sci. Comput. 41 (2019)):
 *
 *   c_j = sum_l b_l * phi(X1_j - l1) phi(X2_j - l2) phi(X3_j - l3)
 *
 * where b is the (complex) fine grid of size n^3 produced by the caller
 * (deconvolve + zero-pad + FFT, done in oracle/es_nufft.py with numpy),
 * X_d = x_d * n / (2 pi) folded periodically into [0, n), and
 * phi(z) = exp(beta * (sqrt(1 - (2 z / w)^2) - 1)) for |z| < w/2, 0 outside.
 * The window of a point starts at i0 = ceil(X - w/2) and holds w grid
 * points; grid indices wrap modulo n.  Axis 0 of the grid pairs with x
 * (finufftpy takes the model in Fortran order with x fastest;
 * cufinufft_ext.py:121-122 feeds (L, K, H) to the C-ordered cufinufft for
 * the same pairing).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#define ORACLE_MAXW 16

static inline double es_phi(double z, int w, double beta)
{
    double t = 2.0 * z / (double)w;
    double v = 1.0 - t * t;
    if (v <= 0.0) return 0.0;
    return exp(beta * (sqrt(v) - 1.0));
}

/* fold x (radians, |x| < 3 pi expected) into grid units in [0, n) */
static inline double fold_rescale(double x, int64_t n)
{
    const double two_pi = 6.283185307179586476925286766559;
    double X = x / two_pi;
    X -= floor(X);
    X *= (double)n;
    if (X >= (double)n) X -= (double)n;
    return X;
}

/*
 * grid: complex128 interleaved, C order [l1][l2][l3], n^3 entries.
 * x, y, z: N coordinates.  out: complex128 interleaved, N entries.
 * Returns 0 on success, -1 for bad arguments.
 */
int oracle_es_interp3d(const double *grid, int64_t n, const double *x,
                       const double *y, const double *z, int64_t N, int w,
                       double beta, double *out)
{
    if (w < 2 || w > ORACLE_MAXW || n < 2 * w || N < 0) return -1;
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < N; ++j) {
        double X[3] = {fold_rescale(x[j], n), fold_rescale(y[j], n),
                       fold_rescale(z[j], n)};
        int64_t i0[3];
        double ker[3][ORACLE_MAXW];
        for (int d = 0; d < 3; ++d) {
            i0[d] = (int64_t)ceil(X[d] - 0.5 * w);
            for (int a = 0; a < w; ++a)
                ker[d][a] = es_phi((double)(i0[d] + a) - X[d], w, beta);
        }
        double re = 0.0, im = 0.0;
        for (int a = 0; a < w; ++a) {
            int64_t l1 = ((i0[0] + a) % n + n) % n;
            for (int b = 0; b < w; ++b) {
                int64_t l2 = ((i0[1] + b) % n + n) % n;
                double wab = ker[0][a] * ker[1][b];
                const double *row = grid + 2 * ((l1 * n + l2) * n);
                double rr = 0.0, ri = 0.0;
                for (int c = 0; c < w; ++c) {
                    int64_t l3 = ((i0[2] + c) % n + n) % n;
                    rr += ker[2][c] * row[2 * l3];
                    ri += ker[2][c] * row[2 * l3 + 1];
                }
                re += wab * rr;
                im += wab * ri;
            }
        }
        out[2 * j] = re;
        out[2 * j + 1] = im;
    }
    return 0;
}

/*
 * Brute-force FP64 direct non-uniform DFT, the ground truth the NUFFT is
 * checked against:  c_j = sum_k f_k exp(isign * i * k . x_j),
 * k_d in [-(M/2), (M-1)/2] (finufft CMCL mode order: f index i <-> k = i - M/2).
 * f: real FP64, C order [k1][k2][k3], M^3 entries.  out: complex interleaved.
 */
int oracle_nudft3d2_real(const double *f, int64_t M, const double *x,
                         const double *y, const double *z, int64_t N,
                         int isign, double *out)
{
    if (M < 1 || N < 0) return -1;
    int64_t h = M / 2;
#pragma omp parallel
    {
        double *c1 = (double *)malloc(sizeof(double) * 6 * M);
        double *s1 = c1 + M, *c2 = c1 + 2 * M, *s2 = c1 + 3 * M;
        double *c3 = c1 + 4 * M, *s3 = c1 + 5 * M;
#pragma omp for schedule(static)
        for (int64_t j = 0; j < N; ++j) {
            for (int64_t i = 0; i < M; ++i) {
                double k = (double)(i - h);
                c1[i] = cos(k * x[j]); s1[i] = isign * sin(k * x[j]);
                c2[i] = cos(k * y[j]); s2[i] = isign * sin(k * y[j]);
                c3[i] = cos(k * z[j]); s3[i] = isign * sin(k * z[j]);
            }
            double re = 0.0, im = 0.0;
            for (int64_t a = 0; a < M; ++a) {
                double r2 = 0.0, i2 = 0.0;
                for (int64_t b = 0; b < M; ++b) {
                    const double *fr = f + (a * M + b) * M;
                    double r3 = 0.0, i3 = 0.0;
                    for (int64_t c = 0; c < M; ++c) {
                        r3 += fr[c] * c3[c];
                        i3 += fr[c] * s3[c];
                    }
                    /* (r3 + i i3) * (c2 + i s2) */
                    r2 += r3 * c2[b] - i3 * s2[b];
                    i2 += r3 * s2[b] + i3 * c2[b];
                }
                re += r2 * c1[a] - i2 * s1[a];
                im += r2 * s1[a] + i2 * c1[a];
            }
            out[2 * j] = re;
            out[2 * j + 1] = im;
        }
        free(c1);
    }
    return 0;
}
