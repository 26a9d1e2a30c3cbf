/*
 * libmtip -- C ABI of the MI355X (gfx950) orientation-matching path:
 * type-2 3D NUFFT slicing of an autocorrelation model at rotated
 * Ewald-sphere pixel q-vectors, slice-to-image distances with the cross
 * term on MFMA, and per-image argmin / top-K with FP32/FP64 rescoring.
 *
 * Conventions
 *  - Every pointer named as "device" is device memory on the plan's /
 *    call's device; the library never copies between host and device
 *    except for its own small tables.  Host pointers are marked "host".
 *  - Every call is asynchronous on `stream` (a hipStream_t; NULL = the
 *    null stream) unless stated otherwise.
 *  - Every function returns MTIP_OK (0) or a negative MTIP_ERR_* code;
 *    mtip_last_error() then gives a thread-local message.
 *  - Arrays are C order.  Models are cubic M^3 float32, model index i along
 *    an axis is frequency k = i - M/2 (finufft's CMCL mode order); axis 0
 *    pairs with H (x), axis 2 with L (z).
 *  - No entry point allocates device memory except mtip_plan_create /
 *    mtip_plan_set_model (the plan's fine grid and tables) and the first
 *    mtip_plan_spread_begin on a plan (its spreading grid); workspaces are
 *    passed in by the caller with their size queried first.
 */
#ifndef MTIP_H
#define MTIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MTIP_ABI_VERSION 5

enum {
    MTIP_OK = 0,
    MTIP_ERR_ARG = -1,          /* bad shape, size, pointer or option */
    MTIP_ERR_DOMAIN = -2,       /* NU point outside [-3 pi, 3 pi) */
    MTIP_ERR_HIP = -3,          /* a HIP runtime call failed */
    MTIP_ERR_NOMEM = -4,        /* device allocation failed / workspace too small */
    MTIP_ERR_STATE = -5,        /* call order (e.g. execute before set_model) */
    MTIP_ERR_UNSUPPORTED = -6   /* option not implemented */
};

/* Arithmetic of the cross-term GEMM. */
enum { MTIP_FP32 = 0, MTIP_BF16 = 1, MTIP_FP8 = 2 };

int mtip_abi_version(void);
const char *mtip_last_error(void);
/* host out: free and total bytes of HBM on `device` (admission control). */
int mtip_mem_info(int device, size_t *free_bytes, size_t *total_bytes);

/* ------------------------------------------------------------------ NUFFT */

typedef struct mtip_plan mtip_plan;

/*
 * Type-2 3D NUFFT plan for a cubic M^3 real model at tolerance eps
 * (ES kernel, upsampling 2, width ceil(-log10(eps/10))).
 * Replaces cufinufft(2, ugrid.shape, 1, eps, isign, dtype, gpu_method=1,
 * gpu_device_id) -- spinifel/extern/nufft_ext.py:271-280,
 * spinifel/extern/cufinufft_ext.py:340-342 -- and the setup half of
 * finufftpy.nufft3d2 (spinifel/extern/finufft_ext.py:88).
 * The plan computes the real part of the transform only, which is all the
 * reference keeps (spinifel/sequential/orientation_matching.py:241,
 * spinifel/slicing.py:43); the sign of the exponent therefore does not
 * matter for it.
 */
int mtip_plan_create(int M, double eps, int device, mtip_plan **out);
int mtip_plan_destroy(mtip_plan *plan);
/* host out: fine grid n, kernel width w, ES beta, padded grid edge. */
int mtip_plan_info(const mtip_plan *plan, int *n_fine, int *width, double *beta,
                   int *padded_edge);

/*
 * Load a model: symmetrise (Re of the transform = transform of
 * (a(k) + a(-k)) / 2), deconvolve by the kernel's Fourier series, and
 * evaluate the real padded fine grid by one FP64 n^3 FFT (hipFFT; the first
 * call on a plan allocates its buffer and plan); the grid stays resident in
 * the plan.  ac: device float32 M^3.
 * Replaces the ugrid upload and the deconvolve/FFT inside
 * plan.execute(nuvect, ugrid) (spinifel/extern/nufft_ext.py:283,
 * cufinufft_ext.py:344) and autocorrelation.forward's `ugrid *= support`
 * with support = 1 (spinifel/autocorrelation.py:53-54).
 */
int mtip_plan_set_model(mtip_plan *plan, const float *ac, void *stream);

/*
 * Re of c_j = sum_k f_k exp(-i k . (x_j, y_j, z_j)) at N arbitrary points
 * (radians).  x, y, z: device FP64; out: device FP32 (N).  Points with a
 * coordinate outside [-3 pi, 3 pi) (or not finite) give NaN and add 1 to
 * *bad_count (device int32, nullable).
 * Replaces plan.set_pts(H_, K_, L_) + plan.execute(nuvect, ugrid)
 * (spinifel/extern/nufft_ext.py:282-283, cufinufft_ext.py:343-344) and
 * finufftpy.nufft3d2(x, y, z, c, isign, eps, f) (finufft_ext.py:88).
 */
int mtip_plan_execute(mtip_plan *plan, const double *x, const double *y, const double *z,
                      int64_t N, float *out, int *bad_count, void *stream);

/*
 * Fused rotation + type-2 slicing for N_o orientations x N_p pixels.
 * For each orientation n (quaternion [w, x, y, z], device FP64 (N_o, 4))
 * the library forms R(q) by skopi's formula (eval/geom.py:26-46) and its
 * FP64 inverse (spinifel/sequential/orientation_matching.py:220), rotates
 * the pixel q-vectors q (device FP64 (3, N_p)) as
 * einsum("ijk,klmn->jilmn") does (:233), scales by `scale` = pi / (E *
 * oversampling) (:234-236), and evaluates Re of the type-2 transform.
 * out: device FP32 (N_o, ld_out), orientation-major like model_slices_new
 * (:238-241).  Optional (nullable) outputs: out_bf16 (N_o, ld_out) bf16 bit
 * patterns of the same values; row_sum / row_sumsq: device FP64 (N_o)
 * sums of s and s^2 per orientation (for the std scaling :245-248 and the
 * squared norms of the distance expansion).  bad_count as for execute.
 * pixel_order (nullable, device int32, n_order entries, n_order a multiple
 * of tile_points): the order in which pixels are visited, each run of
 * tile_points entries one tile of neighbouring pixels (16 x 16, 8 x 16 or
 * 8 x 8 detector blocks for tile_points 256, 128, 64; 0 means 256), entries
 * -1 for padding; every pixel in [0, N_p) must appear exactly once.  Tiles
 * of neighbouring pixels let the kernel stage one small box of the fine grid
 * per tile in LDS; smaller tiles keep the box of a large grid (M >= 121)
 * within LDS.  NULL visits pixels in index order, tile_points at a time.
 * orient_order (nullable, device int32, a permutation of [0, N_o)): the
 * order in which orientations are processed (e.g. a space-filling-curve
 * order, so that concurrently resident workgroups share L2); rows of `out`
 * stay indexed by orientation either way.
 * Replaces the batched loop of spinifel/sequential/orientation_matching.py:230-241
 * and gen_model_slices (spinifel/slicing.py:10-45).
 */
int mtip_plan_slices(mtip_plan *plan, const double *quats, int64_t N_o, const double *q,
                     int64_t N_p, const int32_t *pixel_order, int64_t n_order, int tile_points,
                     const int32_t *orient_order, double scale, float *out, int64_t ld_out,
                     uint16_t *out_bf16, double *row_sum, double *row_sumsq, int *bad_count,
                     void *stream);

/*
 * Rotated, normalised q-points only: H/K/L (device FP64, N_o * N_p,
 * orientation-major) = scale * inv(R(q_n)) . q_j with R from skopi's
 * formula, as gen_nonuniform_normalized_positions computes on the host
 * (spinifel/autocorrelation.py:290-323; orientation_matching.py:220, :233-236).
 * N_o <= 65535 per call.
 */
int mtip_rotate_positions(const double *quats, int64_t N_o, const double *q, int64_t N_p,
                          double scale, double *H, double *K, double *L, void *stream);

/* ------------------------------------------------ type-1 NUFFT (adjoint) */

/*
 * The adjoint of slicing, in three steps on a plan whose M is the output
 * edge (M for the data term, M_ups = 2M for the convolution kernel):
 *
 *   out[i] = scale * support[i] * Re sum_j c_j exp(+i k(i) . x_j),
 *   k_d(i) = i_d - M/2,  out (M, M, M) C order, axis 0 paired with x (= H)
 *
 * i.e. spinifel's adjoint (spinifel/autocorrelation.py:97-128 and
 * spinifel/extern/nufft_ext.py:287-327: nufft_3d_t1 with isign = +1, times
 * support, real part, / M^3 when scale = 1 / M^3), evaluated by the plan's
 * ES kernel: spreading onto the fine grid, then the transform back to modes.
 *
 * mtip_plan_spread_begin clears the plan's FP64 spreading grid (allocating
 * it and the pass buffers on first use: (np^3 + 3 np^2 (M/2 + 1)) * 8 bytes
 * roughly).  Any number of spread calls then accumulate; spread_finish runs
 * the transform and writes `out` (device FP64, M^3); the grid keeps its
 * contents until the next begin.  Calls on one plan must not overlap.
 */
int mtip_plan_spread_begin(mtip_plan *plan, void *stream);

/*
 * Spread row n of `values` (device FP32 (N_o, ld_values), indexed by pixel;
 * NULL spreads ones, as Merge's nuvect of ones, sequential/autocorrelation.py:70)
 * at the q-points of orientation n: scale * inv(R(quats[n])) . q_j, the
 * positions of Merge.get_non_uniform_positions
 * (spinifel/sequential/autocorrelation.py:85-108) times mult (:146-148).
 * quats, q, pixel_order, n_order, tile_points as for mtip_plan_slices (a
 * tile's points share one LDS box: 8 x 8 tiles keep the boxes of fine
 * grids above ~200 points per edge within LDS).  bad_count (nullable device int): += points outside |x| < 3 pi
 * (skipped; spinifel/autocorrelation.py:114 asserts instead).
 */
int mtip_plan_spread_slices(mtip_plan *plan, const double *quats, int64_t N_o, const double *q,
                            int64_t N_p, const int32_t *pixel_order, int64_t n_order,
                            int tile_points, double scale, const float *values,
                            int64_t ld_values, int *bad_count, void *stream);

/*
 * Spread values (device FP32 (N), NULL = ones) at arbitrary points
 * (x, y, z device FP64 (N), radians): nufft_3d_t1(H_, K_, L_, nuvect, ...)
 * of spinifel/autocorrelation.py:119 / extern/nufft_ext.py:316.  Tiles of 128
 * consecutive points share an LDS box, so neighbouring points should be
 * adjacent (as a flattened detector is, or points put in a spatial order).
 */
int mtip_plan_spread_points(mtip_plan *plan, const double *x, const double *y, const double *z,
                            int64_t N, const float *values, int *bad_count, void *stream);

/*
 * Transform the accumulated grid to the M^3 real output (see above);
 * support: device FP64 (M^3) or NULL for 1.
 */
int mtip_plan_spread_finish(mtip_plan *plan, const double *support, double scale, double *out,
                            void *stream);

/* ---------------------------------------------------------------- phasing */

/*
 * The two projections of each ER / HIO iteration of spinifel's phasing
 * (spinifel/sequential/phasing.py:129-194), FP64, for a real density whose
 * spectrum the caller keeps as the half spectrum of a real-to-complex
 * transform (Hermitian, so the replaced spectrum's inverse is the real part
 * the reference keeps).
 *
 * mtip_phase_amplitudes: spectrum (device complex128, n entries, in place)
 * <- amplitudes * exp(i angle(spectrum)) everywhere but entry 0, the masked
 * zero frequency (step_phase, :187-190, with amp_mask_ of :269-270).
 *
 * mtip_phase_update: rho (device FP64, n, in place) from the projected
 * density rho_mod and the support (device uint8): ER (mode 0, :141-146) or
 * HIO with feedback beta (mode 1, :162-167), with the rho_max clamp
 * (+inf as phase() passes, :279).
 */
int mtip_phase_amplitudes(void *spectrum, const double *amplitudes, int64_t n, void *stream);
int mtip_phase_update(double *rho, const double *rho_mod, const uint8_t *support, int64_t n,
                      int mode, double beta, double rho_max, void *stream);

/* ---------------------------------------------------------------- merge */

/*
 * Vector work of the merge's CG iterations (FP64, device pointers), around
 * the caller's real-to-complex FFTs of
 *   W u = s . crop(irfft(rfft(pad(s . u)) . F)) + rlambda u + flambda s . irfft(rfft(s . u) . mask)
 * (spinifel/sequential/autocorrelation.py:110-145, :173-178).
 * mtip_merge_pad: ups[:M, :M, :M] = s . u inside the caller-zeroed
 * (M_ups)^3 array (the rest is left alone) and us = s . u (M^3);
 * support may be NULL for 1.
 * mtip_merge_combine: out = s . y_ups[:M, :M, :M] + rlambda u + flambda s . y_reg.
 */
int mtip_merge_pad(const double *u, const double *support, int M, int M_ups, double *ups,
                   double *us, void *stream);
int mtip_merge_combine(const double *y_ups, const double *y_reg, const double *u,
                       const double *support, int M, int M_ups, double rlambda, double flambda,
                       double *out, void *stream);

/*
 * Conjugate-gradient steps of cupyx/scipy cg (:211-212) with the step
 * lengths on the device: mtip_dot writes a . b to *out (two fixed-order
 * reduction passes, workspace of mtip_dot_workspace_size() bytes);
 * mtip_cg_update: x += (rho / pq) p, r -= (rho / pq) q;
 * mtip_cg_direction: p = r (first) or r + (rho / rho_old) p.
 * rho, rho_old, pq: device FP64 scalars.
 */
size_t mtip_dot_workspace_size(void);
int mtip_dot(const double *a, const double *b, int64_t n, double *work, size_t work_bytes,
             double *out, void *stream);
int mtip_cg_update(double *x, double *r, const double *p, const double *q, const double *rho,
                   const double *pq, int64_t n, void *stream);
int mtip_cg_direction(double *p, const double *r, const double *rho, const double *rho_old,
                      int64_t n, int first, void *stream);

/* ------------------------------------------------------------- data prep */

/*
 * Clipping to the low-resolution region followed by 2^n_bin x 2^n_bin
 * binning, in one pass: spinifel/prep.py's binning_sum / binning_mean /
 * binning_index = bin2nx2n_{sum,mean,index}(clipping(arr, N_clipping),
 * N_binning) (prep.py:36-134), with the reductions in the reference's order
 * and type.  in: device (n_lead, panels, sa, sb) C order of dtype 0 float32,
 * 1 float64, 2 int32, 3 int64; out: device (n_lead, panels, sa >> (n_clip +
 * n_bin), sb >> (n_clip + n_bin)) of the same type.  mode 0 sum, 1 mean
 * (float only), 2 index (min of each 2 x 2 block, floor-halved, per level).
 * Clipping keeps the centred window with 1 panel and the bottom corners with
 * 4 panels (prep.py:86-100); other panel counts give zeros, as the
 * reference's array stays zero.  The clipped panel edges must be multiples of
 * 2^n_bin (n_bin <= 4).  clipping_index's subtraction of the minimum
 * (:108-113) is left to the caller.
 */
int mtip_clip_bin(const void *in, int dtype, int64_t n_lead, int panels, int sa, int sb,
                  int n_clip, int n_bin, int mode, void *out, void *stream);

/* ---------------------------------------------------------- row utilities */

/* Per-row FP64 sum and sum of squares of a float32 (rows, cols) matrix. */
int mtip_row_stats(const float *in, int64_t rows, int64_t cols, int64_t ld, double *row_sum,
                   double *row_sumsq, void *stream);

/*
 * alpha = std(images) / std(slices), population std over all entries, FP64,
 * from per-row sums (spinifel/sequential/orientation_matching.py:246).
 * alpha: device FP64 scalar (stays on the device for mtip_match_topk).
 */
int mtip_scaling_ratio(const double *img_sum, const double *img_sumsq, int64_t n_img,
                       const double *sl_sum, const double *sl_sumsq, int64_t n_sl,
                       int64_t n_pix, double *alpha, void *stream);

/*
 * Per-request scaling ratio: for each segment g of images [seg_offsets[g],
 * seg_offsets[g+1]) (device int64, n_seg + 1 entries) alpha[i] = std(images
 * of the segment) / std(slices) for every image i of it (device FP64, N_i).
 * A batch of independent requests then matches each request exactly as if
 * it were matched alone.
 */
int mtip_segment_scaling_ratio(const double *img_sum, const double *img_sumsq,
                               const int64_t *seg_offsets, int64_t n_seg, const double *sl_sum,
                               const double *sl_sumsq, int64_t n_sl, int64_t n_pix, double *alpha,
                               void *stream);

/*
 * Cast rows to the GEMM operand format.  format MTIP_BF16: out is uint16
 * bf16 (round to nearest even), scales unused (nullable).  format MTIP_FP8:
 * out is uint8 OCP e4m3 with a per-row scale = absmax / 448 written to
 * scales (device float32 (rows)), value = e4m3 * scale.
 */
int mtip_quantize_rows(const float *in, int64_t rows, int64_t cols, int64_t ld_in, int format,
                       void *out, int64_t ld_out, float *scales, void *stream);

/*
 * Residual rows for the GEMM operands: out_r = cast(m_r * in_r - center) with
 * m_r from the device FP64 array mult: mult[0] for every row, or mult[r]
 * when bit 1 of mult_flags is set; its inverse when bit 0 is set; 1 when
 * mult is NULL.  center a
 * device float32 (cols) vector (NULL = 0), format MTIP_FP32 / MTIP_BF16 /
 * MTIP_FP8 (e4m3 with per-row scales absmax / 448).  sumsq (nullable):
 * device FP64 (rows) squared norms of the values as cast.  Distances are
 * invariant to subtracting one vector from both slices and images, so a
 * common centre removes the orientation-independent part of the pattern
 * (the bright low-q region) before the cast.  Columns [cols, ld_out) are 0.
 */
int mtip_residual_rows(const float *in, int64_t rows, int64_t cols, int64_t ld_in,
                       const double *mult, int mult_flags, const float *center, int format,
                       void *out, int64_t ld_out, float *scales, double *sumsq, void *stream);

/*
 * Column means (FP64 accumulation, fixed summation order) of a float32
 * (rows, cols) matrix -> float32 (cols).  work: device scratch of
 * mtip_col_mean_workspace_size(cols) bytes.
 */
size_t mtip_col_mean_workspace_size(int64_t cols);
int mtip_col_mean(const float *in, int64_t rows, int64_t cols, int64_t ld, double *work,
                  size_t work_bytes, float *mean, void *stream);

/* ------------------------------------------------------------- matching */

/* host: bytes of workspace mtip_match_topk needs. */
size_t mtip_match_workspace_size(int64_t N_o, int64_t N_i, int K, int precision);

/*
 * host: orientation rows per GEMM tile for `precision`: 0 restores the
 * built-in rule (256 for BF16 and FP8 when the problem has at least 192
 * tiles of 256 x 256, else 128), 128
 * or 256 force a tile.  Process-wide; mtip_match_topk reads it once per
 * call, and mtip_match_partial / mtip_topk_merge take the tile explicitly
 * (tile_rows) so that a change between them cannot split a call.  The
 * 256-row tile needs
 * 256 * ld * element bytes < 2^31 (mtip_match_partial reports
 * MTIP_ERR_UNSUPPORTED otherwise).  No counterpart in the reference (the
 * reference has one GEMM path); the project's autotuner sets it from timings.
 */
int mtip_set_gemm_tile(int precision, int rows);

/* host: the tile rows mtip_match_partial will use for these sizes (>0), or an error code. */
int mtip_gemm_tile(int precision, int64_t N_o, int64_t N_i);

/*
 * Per image i the K orientations with the smallest
 *   v(n, i) = alpha^2 ||s_n||^2 - 2 alpha s_n . d_i
 * (= squared distance minus the image's own ||d_i||^2, which does not move
 * the argmin), cross term on MFMA in `precision` with FP32 accumulation.
 * slices (N_o, ld_s) and images (N_i, ld_d) are float32 / bf16 / e4m3 per
 * `precision`; slice_scale / image_scale per-row dequantisation factors
 * (FP8 only, else nullable); slice_sumsq: device FP64 (N_o) ||s_n||^2 of the
 * unscaled FP32 slices; alpha: device FP64 scalar.  K in {1, 2, 4, 8}.
 * Outputs: top_idx device int32 (N_i, K) and top_val device float32
 * (N_i, K), ascending, lowest index first on equal values (the first-
 * occurrence rule of numpy's argmin, spinifel/sequential/nearest_neighbor.py:135,
 * and the strict < of orientation_matching_sp.cu:549-557).
 * N_p must be a multiple of 64 (32 for FP32); ld_s, ld_d likewise.
 * Replaces SNM.euclidean_dist + dist.argmin (spinifel/sequential/
 * orientation_matching.py:81-109, :179) and the KNN pair
 * cudaEuclideanDistance / cudaHeapSort (orientation_matching_sp.cu:612, :682)
 * without materialising the (N_o, N_i) matrix.
 */
int mtip_match_topk(int precision, const void *slices, int64_t ld_s, const float *slice_scale,
                    const double *slice_sumsq, int64_t N_o, const void *images, int64_t ld_d,
                    const float *image_scale, int64_t N_i, int64_t N_p, const double *alpha,
                    int K, void *workspace, size_t workspace_bytes, int32_t *top_idx,
                    float *top_val, void *stream);

/*
 * The two halves of mtip_match_topk, for callers that time or reuse them:
 * mtip_match_partial runs the MFMA GEMM and writes per-(image, orientation
 * tile) top-K packed keys to the workspace; mtip_topk_merge reduces them to
 * the per-image top-K and must get the same precision, sizes and tile_rows.
 * tile_rows: 128 or 256 orientation rows per GEMM tile, or 0 for the value
 * mtip_gemm_tile gives at the time of the call.  Pass the value
 * mtip_gemm_tile returned to both calls when another thread may call
 * mtip_set_gemm_tile in between (0 would then read the new setting in one
 * call and the old one in the other).  Other arguments and rules as
 * mtip_match_topk.
 */
int mtip_match_partial(int precision, const void *slices, int64_t ld_s, const float *slice_scale,
                       const double *slice_sumsq, int64_t N_o, const void *images, int64_t ld_d,
                       const float *image_scale, int64_t N_i, int64_t N_p, const double *alpha,
                       int K, int tile_rows, void *workspace, size_t workspace_bytes,
                       void *stream);
int mtip_topk_merge(int precision, const void *workspace, int64_t N_o, int64_t N_i, int K,
                    int tile_rows, int32_t *top_idx, float *top_val, void *stream);

/*
 * mtip_match_partial with flags.  MTIP_MATCH_SPLIT_EDGES: on the 256-row
 * tile, the thin edge tiles (a last row of orientation tiles or column of
 * image tiles with at most 64 valid rows) run as a short second launch, each
 * split in parts of K whose FP32 sums are added in a fixed order, after the
 * whole tiles: a thin tile takes as long as a whole one, so without the split
 * they can add a nearly empty round of tiles (10,000 x 10,000: 1,600 tiles on
 * 256 CUs, 79 of them thin).  The split applies when its partial sums fit in
 * mtip_match_workspace_size(N_o, N_i, K, precision) bytes beside the keys, so
 * whether it applies depends on the sizes only; the workspace must then hold
 * that size.  Keys of the split tiles can differ from mtip_match_partial's in
 * the last bits (another summation order); pass 0 for exactly its results
 * (callers that need one summation order whatever the batch, e.g. serving).
 * No counterpart in the reference (one GEMM path there).
 */
#define MTIP_MATCH_SPLIT_EDGES 1
int mtip_match_partial_ex(int precision, const void *slices, int64_t ld_s,
                          const float *slice_scale, const double *slice_sumsq, int64_t N_o,
                          const void *images, int64_t ld_d, const float *image_scale, int64_t N_i,
                          int64_t N_p, const double *alpha, int K, int tile_rows, int flags,
                          void *workspace, size_t workspace_bytes, void *stream);

/*
 * Bytes of the device scratch mtip_match_topk_static uses (32 MiB), one
 * copy per GPU the library runs on, held until the process ends.
 */
#define MTIP_STATIC_SCRATCH_BYTES (32u << 20)

/*
 * mtip_match_topk without a workspace argument: the top-K partial keys go
 * to a device array inside the library (MTIP_STATIC_SCRATCH_BYTES per GPU),
 * so the call allocates no device memory and calls no allocating routine.
 * No counterpart in the reference (its KNN allocates and frees device
 * buffers on every call, orientation_matching_sp.cu:625-627, :677); it is
 * for callers that cannot allocate: code under graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.
 * Limit: N_i * ceil(N_o / tile) * K * 8 bytes <= MTIP_STATIC_SCRATCH_BYTES
 * with tile = mtip_gemm_tile(precision, N_o, N_i); larger calls return
 * MTIP_ERR_ARG (split the images into several calls).  Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU (other streams
 * or threads) race and must be serialised by the caller.  Same arguments
 * and results as mtip_match_topk otherwise.
 */
int mtip_match_topk_static(int precision, const void *slices, int64_t ld_s,
                           const float *slice_scale, const double *slice_sumsq, int64_t N_o,
                           const void *images, int64_t ld_d, const float *image_scale,
                           int64_t N_i, int64_t N_p, const double *alpha, int K,
                           int32_t *top_idx, float *top_val, void *stream);

/*
 * Rescore K candidates per image on FP32 slices/images:
 *   d(n, i) = sum_p (alpha s_n[p] - d_i[p])^2
 * with alpha[0] for every image, or alpha[i] per image when
 * alpha_per_image (a request's own scaling ratio, so that its result does not
 * depend on which other images share the batch),
 * once with FP32 arithmetic (d32, nullable) and once in FP64 (d64,
 * nullable); best_idx / best_d64 (device, N_i) take the FP64 minimum
 * (lowest orientation index on ties).  cand: device int32 (N_i, K);
 * entries outside [0, N_o) are skipped (their d32 / d64 are NaN).  This is the distance of
 * spinifel/sequential/nearest_neighbor.py:134 (squared, before the sqrt).
 * image_order (nullable, device int32, a permutation of [0, N_i)): the order
 * in which images are processed (images with shared candidates next to each
 * other reuse the candidates' slices from cache); outputs stay per image.
 */
int mtip_rescore(const float *slices, int64_t ld_s, int64_t N_o, const float *images, int64_t ld_d,
                 int64_t N_i, int64_t N_p, const double *alpha, int alpha_per_image,
                 const int32_t *cand, int K,
                 int32_t *best_idx, double *best_d64, float *d32, double *d64,
                 const int32_t *image_order, void *stream);

/*
 * Full squared-distance matrix euDist[i * numRef + n] = ||data_i - ref_n||^2
 * (FP32 MFMA cross term, FP64-accumulated norms), all totalPixels pixels.
 * Replaces cudaEuclideanDistance(refImgs, dataImgs, numRefImgs,
 * numDataImgs, totalPixels, deviceId) (orientation_matching_sp.cu:612-679)
 * minus its host copies and its dropping of the last totalPixels % 16
 * pixels (:110).  totalPixels must be a multiple of 4 (rows are read in
 * 16-byte chunks); ref and data 16-byte aligned.
 */
int mtip_euclidean_distance(const float *ref, const float *data, int64_t numRef, int64_t numData,
                            int64_t totalPixels, float *euDist, void *stream);

/*
 * Indices of the totalNN smallest entries of each row of euDist
 * (numData, numRef), ascending, lowest index first on ties: idx device
 * int32 (numData, totalNN); entries past numRef are -1.  Any totalNN >= 1:
 * up to 8 by one register top-K pass per row, beyond that by a radix
 * select and an LDS sort per row, 2048 indices at a time.
 * Replaces cudaHeapSort(euDist, numDataImgs, numRefImgs, totalPixels,
 * totalNN, deviceId) (orientation_matching_sp.cu:682-729).
 */
int mtip_heap_sort(const float *euDist, int64_t numData, int64_t numRef, int totalNN,
                   int32_t *idx, void *stream);

/* ------------------------------------------------------------- orientation hierarchy */

/*
 * Batched refinement over a hierarchy of orientation grids (no counterpart
 * in the reference, whose matching is one flat grid: the batched loop of
 * spinifel/sequential/orientation_matching.py:230-256 over all N_o
 * orientations).  After a dense top-T against a coarse grid, each image
 * refines among the children of its T winners on the next grid.  An entry
 * is one (image i, winner t) pair, e = i * n_win + t.
 *
 * host: bytes of the grouping workspace for N_e entries and n_groups coarse nodes.
 */
size_t mtip_group_workspace_size(int64_t N_e, int64_t n_groups);

/*
 * Groups entries by winner: winners device int32, winner t of image i at
 * winners[i * win_stride + t] (t < n_win <= win_stride; e.g. the first
 * n_win columns of a (N_i, K) top-K index), in [0, n_groups) (others are
 * counted as bad and left out).  Fills the workspace with group sizes and
 * offsets, the entries of each group and a task list of (group, up to 32
 * entries), all on the device: no host synchronisation and no memset, so
 * the call can be captured in a graph.
 */
int mtip_group_by_winner(const int32_t *winners, int64_t win_stride, int n_win, int64_t N_i,
                         int64_t n_groups, void *workspace, size_t ws_bytes, void *stream);

/*
 * Per entry, the K children of its winner with the smallest
 *   v = ||s_c||^2 - 2 s_c . d_i
 * on bf16 MFMA with FP32 accumulation (the centred, scaled operands of
 * mtip_match_topk with alpha = 1).  child_off (n_groups + 1) / child_idx:
 * device int32 CSR of each coarse node's children (row indices of
 * `slices`); slices (N_fine, ld_s) bf16, slice_sumsq FP64 (N_fine);
 * images (N_i, ld_d) bf16; kd: pixels summed, a multiple of 64 within both
 * pitches (columns past N_p zero).  A persistent kernel walks the task list
 * the previous mtip_group_by_winner (same n_win, N_i, n_groups) left in
 * `workspace`.  Outputs top_idx int32 / top_val float32 (N_i * n_win, K),
 * row e for entry e, ascending, lowest index first on ties, -1 / +inf past
 * the group's children; entries left out are not written.  K in {1, 2, 4, 8}.
 */
int mtip_group_topk(const void *workspace, int n_win, int64_t N_i, int64_t n_groups,
                    const int32_t *child_off, const int32_t *child_idx, const uint16_t *slices,
                    int64_t ld_s, const double *slice_sumsq, const uint16_t *images, int64_t ld_d,
                    int64_t kd, int K, int32_t *top_idx, float *top_val, int device, void *stream);

/*
 * mtip_group_topk for either operand format: precision MTIP_BF16 (as
 * mtip_group_topk; the scales are ignored) or MTIP_FP8 (e4m3 rows with
 * per-row float scales slice_scale (N_fine) and image_scale (N_i), as
 * mtip_quantize_rows / mtip_residual_rows write them; the cross term is
 * scale_c scale_e (q_c . q_e) and slice_sumsq holds the squared norms of the
 * scaled values).  Pitches ld_s / ld_d and kd in elements; rows must be
 * whole 16-byte units.  No counterpart in the reference (batching form).
 */
int mtip_group_topk_prec(int precision, const void *workspace, int n_win, int64_t N_i,
                         int64_t n_groups, const int32_t *child_off, const int32_t *child_idx,
                         const void *slices, int64_t ld_s, const float *slice_scale,
                         const double *slice_sumsq, const void *images, int64_t ld_d,
                         const float *image_scale, int64_t kd, int K, int32_t *top_idx,
                         float *top_val, int device, void *stream);

/*
 * host: byte offset in the grouping workspace of the int32 entry permutation
 * (N_e entries: each group's entries together, groups in order, entries
 * with no valid winner last), valid after mtip_group_by_winner; usable as
 * the image_order of mtip_rescore when n_win = 1.
 */
size_t mtip_group_perm_offset(int64_t N_e, int64_t n_groups);

/* Number of entries the last mtip_group_by_winner left out (device int32 copy). */
int mtip_group_bad_count(const void *workspace, int n_win, int64_t N_i, int64_t n_groups,
                         int32_t *out, void *stream);

/*
 * In place, every entry of a row of idx (device int32 (rows, cols)) equal to
 * an earlier entry of the same row becomes -1, so that a later
 * mtip_group_by_winner leaves the repeat out (no counterpart in the
 * reference: the batching form keeps each entry's best children as the
 * next level's winners, and two entries can reach the same child).
 */
int mtip_unique_rows(int32_t *idx, int64_t rows, int cols, void *stream);

/*
 * Per image, the K_out smallest distinct (value, index) candidates among
 * its n_in (in_idx / in_val, (N_i, n_in), index -1 = none): merges an
 * image's entries after mtip_group_topk.  Outputs (N_i, K_out) ascending,
 * lowest index first on ties, -1 / +inf when fewer.  K_out in {1, 2, 4, 8}.
 */
int mtip_merge_distinct(const int32_t *in_idx, const float *in_val, int64_t N_i, int n_in,
                        int K_out, int32_t *out_idx, float *out_val, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MTIP_H */
