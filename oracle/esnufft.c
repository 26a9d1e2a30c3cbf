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
#ifdef _OPENMP
#include <omp.h>
#endif

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

/*
 * Spreading step of the ES type-1 3D NUFFT that spinifel calls as
 * nufft_3d_t1(H_, K_, L_, nuvect, +1, eps, M, M, M)
 * (spinifel/autocorrelation.py:119, sequential/autocorrelation.py:158-166 via
 * extern/nufft_ext.py:287-327; finufftpy nufft3d1 at
 * extern/finufft_ext.py:41-64):
 *
 *   b_l += c_j * phi(l1 - X1_j) phi(l2 - X2_j) phi(l3 - X3_j)
 *
 * on the periodic grid of n^3 points, X as for oracle_es_interp3d.  The
 * caller (oracle/es_nufft.py) applies the FFT and deconvolution.  Each
 * thread owns a slab of l1 planes and walks every point, so the sums are
 * race-free and deterministic.
 * c: complex interleaved (N).  grid: complex interleaved n^3, accumulated.
 */
int oracle_es_spread3d(const double *x, const double *y, const double *z,
                       const double *c, int64_t N, int64_t n, int w, double beta,
                       double *grid)
{
    if (w < 2 || w > ORACLE_MAXW || n < 2 * w || N < 0) return -1;
#pragma omp parallel
    {
        int nt = 1, t = 0;
#ifdef _OPENMP
        nt = omp_get_num_threads();
        t = omp_get_thread_num();
#endif
        int64_t p0 = n * t / nt, p1 = n * (t + 1) / nt;
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
            double cr = c[2 * j], ci = c[2 * j + 1];
            for (int a = 0; a < w; ++a) {
                int64_t l1 = ((i0[0] + a) % n + n) % n;
                if (l1 < p0 || l1 >= p1) continue;
                for (int b = 0; b < w; ++b) {
                    int64_t l2 = ((i0[1] + b) % n + n) % n;
                    double wab = ker[0][a] * ker[1][b];
                    double *row = grid + 2 * ((l1 * n + l2) * n);
                    for (int e = 0; e < w; ++e) {
                        int64_t l3 = ((i0[2] + e) % n + n) % n;
                        double wt = wab * ker[2][e];
                        row[2 * l3] += wt * cr;
                        row[2 * l3 + 1] += wt * ci;
                    }
                }
            }
        }
    }
    return 0;
}

/*
 * Brute-force FP64 direct type-1 transform, the ground truth of the adjoint:
 *   f_k = sum_j c_j exp(isign * i * k . x_j),  k_d in [-(M/2), (M-1)/2],
 * f index i <-> k = i - M/2 (CMCL), C order [k1][k2][k3], axis 0 with x.
 * c: complex interleaved (N).  out: complex interleaved M^3.
 */
int oracle_nudft3d1(const double *x, const double *y, const double *z,
                    const double *c, int64_t N, int64_t M, int isign, double *out)
{
    if (M < 1 || N < 0) return -1;
    int64_t h = M / 2;
    for (int64_t i = 0; i < 2 * M * M * M; ++i) out[i] = 0.0;
#pragma omp parallel
    {
        double *e = (double *)malloc(sizeof(double) * 6 * M);
        double *c1 = e, *s1 = e + M, *c2 = e + 2 * M, *s2 = e + 3 * M;
        double *c3 = e + 4 * M, *s3 = e + 5 * M;
        /* each thread owns planes k1 in [a0, a1) */
        int nt = 1, t = 0;
#ifdef _OPENMP
        nt = omp_get_num_threads();
        t = omp_get_thread_num();
#endif
        int64_t a0 = M * t / nt, a1 = M * (t + 1) / nt;
        for (int64_t j = 0; j < N; ++j) {
            for (int64_t i = 0; i < M; ++i) {
                double k = (double)(i - h);
                c1[i] = cos(k * x[j]); s1[i] = isign * sin(k * x[j]);
                c2[i] = cos(k * y[j]); s2[i] = isign * sin(k * y[j]);
                c3[i] = cos(k * z[j]); s3[i] = isign * sin(k * z[j]);
            }
            double cr = c[2 * j], ci = c[2 * j + 1];
            for (int64_t a = a0; a < a1; ++a) {
                /* (cr + i ci) * e1 */
                double r1 = cr * c1[a] - ci * s1[a], i1 = cr * s1[a] + ci * c1[a];
                for (int64_t b = 0; b < M; ++b) {
                    double r2 = r1 * c2[b] - i1 * s2[b], i2 = r1 * s2[b] + i1 * c2[b];
                    double *o = out + 2 * ((a * M + b) * M);
                    for (int64_t k3 = 0; k3 < M; ++k3) {
                        o[2 * k3] += r2 * c3[k3] - i2 * s3[k3];
                        o[2 * k3 + 1] += r2 * s3[k3] + i2 * c3[k3];
                    }
                }
            }
        }
        free(e);
    }
    return 0;
}
