/*
 * omxp.h -- C ABI of the omxp MI355X (gfx950) mixed-precision dense solver.
 *
 * The boundary replaces the reference's (at-aaims/OpenMxP) operator layer:
 * its GPU launchers and its vendor-library macros (rocm_device_macros.h).
 * Each entry point names the reference interface it replaces.
 *
 * Conventions
 *  - Column-major storage; all sizes and leading dimensions are int64_t.
 *  - Every buffer is device memory allocated by the caller.  Entry points
 *    never allocate; scratch space is passed in (the *_workspace_bytes
 *    queries give its size).
 *  - FP16 values are IEEE binary16 bit patterns (uint16_t); FP8 values are
 *    OCP e4m3fn bit patterns (uint8_t).
 *  - `stream` is a hipStream_t (NULL = the default stream).  Work is
 *    enqueued asynchronously unless the comment says otherwise.
 *  - Return value: 0 on success, < 0 for an invalid argument or a HIP
 *    runtime error (see OMX_ERR_*), > 0 for numerical information (the
 *    LAPACK-style 1-based index of a zero pivot).
 */
#ifndef OMXP_H
#define OMXP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OMX_OK 0
#define OMX_ERR_ARG (-1)       /* invalid size, pointer or option */
#define OMX_ERR_HIP (-2)       /* a HIP runtime call or launch failed */
#define OMX_ERR_WORKSPACE (-3) /* workspace too small */
#define OMX_ERR_NOCONV (-4)    /* refinement did not converge */

/* Masks of omx_pack_f16, on packed coordinates (r = panel row, k = column
 * within the block): the -alt 2 triangular copies of the inverse factors. */
#define OMX_MASK_NONE 0
#define OMX_MASK_K_LE_R 1      /* keep k <= r, zero elsewhere            */
#define OMX_MASK_K_GT_R_UNIT 2 /* keep k > r, 1 on k == r, zero elsewhere */
#define OMX_MASK_R_GT_K_UNIT 3 /* keep r > k, 1 on r == k, zero elsewhere */

/* Triangular-solve variants of omx_trsm_f32 (T is the triangle). */
#define OMX_TRSM_RUNN 0 /* B <- B * T^-1, T upper, non-unit  */
#define OMX_TRSM_LLNU 1 /* B <- T^-1 * B, T lower, unit      */
#define OMX_TRSM_RLNU 2 /* B <- B * T^-1, T lower, unit      */

/* Library version string (static storage). */
const char *omx_version(void);

/* ---------------------------------------------------------------------------
 * Matrix and right-hand-side generators (64-bit LCG, bit-exact with the
 * reference's hpl_rand.hpp:12-78 and its on-device generator).
 * ------------------------------------------------------------------------- */

/* Fill the whole order-n system matrix A (lda >= n) in FP32 and its FP64
 * diagonal: A_ij = float(LCG value of linear index i + n*j) off the
 * diagonal, diag_i = sum_{j != i} |A_ij| summed in FP64 in the order of the
 * reference's fill_panel_diag_dev (512 strided partial sums, tree to 4,
 * ((s0+s1)+s2)+s3), A_ii = float(diag_i).
 * Replaces panel_matgen_dev (matgen.hpp:84-117) with fill_panel_dev and
 * fill_panel_diag_dev (matgen.hpp:8-82). */
int omx_matgen_f32(int64_t n, uint64_t seed, float *A, int64_t lda,
                   double *diag, void *stream);

/* Raw LCG block A(i0:i0+m, j0:j0+k) of the order-n matrix in FP32 with no
 * diagonal substitution.  Replaces fill_panel_dev (matgen.hpp:67-82) and
 * panel_fill_one_with_rand without calc_diag (hpl_rand.hpp:80-110). */
int omx_matgen_block_f32(int64_t n, uint64_t seed, int64_t i0, int64_t j0,
                         int64_t m, int64_t k, float *A, int64_t lda,
                         void *stream);

/* diag[r] = FP64 off-diagonal row sum of global row i0 + r, r < m, in the
 * reference's on-device order (fill_panel_diag_dev, matgen.hpp:8-65). */
int omx_diag_f64(int64_t n, uint64_t seed, int64_t i0, int64_t m,
                 double *diag, void *stream);

/* b_i = LCG value of linear index n*n + i, i < n (panel_colvgen,
 * matgen.hpp:419-436). */
int omx_rhs_f64(int64_t n, uint64_t seed, double *b, void *stream);

/* ---------------------------------------------------------------------------
 * Refinement kernels (FP64).
 * ------------------------------------------------------------------------- */

/* Bytes of scratch omx_otf_gemv_f64 needs for order n. */
size_t omx_otf_gemv_workspace_bytes(int64_t n);

/* y <- y + alpha * A x with A the generated order-n matrix, regenerated from
 * the LCG on the fly (never read from memory) and A_ii = diag[i].  Each row
 * is summed in FP64 as column segments of fixed length added in order, so
 * results are deterministic.  Replaces otf_gemv_h / otf_gemv_d /
 * otf_gemv_diag_d (panel_gemv.hpp:13-117) and the 1x1 case of panel_gemv_h
 * (panel_gemv.hpp:119-179). */
int omx_otf_gemv_f64(int64_t n, uint64_t seed, const double *diag,
                     double alpha, const double *x, double *y, void *work,
                     size_t work_bytes, void *stream);

/* As omx_otf_gemv_f64 (same kernels, same column segments, bit-identical
 * results) with the partial sums in the library's own static device array
 * of 64 MiB instead of a workspace argument: it allocates nothing, for
 * callers that cannot (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * 64 MiB cover every order n (the largest need, at n = 4194048, is 67104768
 * bytes); larger needs would return OMX_ERR_WORKSPACE.  Every GPU holds its
 * own copy from module load until the process ends; the first call on a
 * device looks up its address (hipGetSymbolAddress), so make one call
 * before any graph capture.  All calls on one GPU share the array: calls
 * that overlap on one GPU (other streams or threads) race.  `stream` must
 * belong to the current device. */
int omx_otf_gemv_f64_noalloc(int64_t n, uint64_t seed, const double *diag,
                             double alpha, const double *x, double *y,
                             void *stream);

/* *result (device) = max_i |x_i|.  Replaces calc_infnorm_d /
 * colv_infnorm_h (panel_norm.hpp:11-58). */
int omx_norm_inf_f64(int64_t n, const double *x, double *result,
                     void *stream);

/* y = x (copycolv_h, panel.hpp:524-545). */
int omx_copy_f64(int64_t n, const double *x, double *y, void *stream);

/* x_i = b_i / d_i (divcolv_h, panel.hpp:547-568). */
int omx_div_f64(int64_t n, const double *b, const double *d, double *x,
                void *stream);

/* y_i = y_i + alpha * x_i (addcolv_h, panel.hpp:603-623; rocblas_daxpy). */
int omx_axpy_f64(int64_t n, double alpha, const double *x, double *y,
                 void *stream);

/* ---------------------------------------------------------------------------
 * Conditioned test matrices ("randsvd"): A = U diag(sigma) V^T with
 * sigma_i = kappa^(-i/(n-1)), U and V products of nrefl seeded Householder
 * reflectors each (reflector t's vector: the LCG values at linear indices
 * n^2 + n + t*n + i, normalised).  Held as A = diag(sigma) + P Q^T with
 * P, Q of n x rank, rank = 2*nrefl <= 16, column-major.  No reference
 * counterpart (the reference's Higham generator, highammgen.hpp:4-30, is a
 * different family and unreachable from its driver).
 * ------------------------------------------------------------------------- */

/* sigma (n), P, Q (n x 2*nrefl) into HOST arrays (FP64). */
int omx_randsvd_factors_host(int64_t n, uint64_t seed, double kappa,
                             int nrefl, double *sigma, double *P, double *Q);

/* As omx_randsvd_factors_host into DEVICE arrays (computed on the host,
 * copied on `stream`; returns after the copies). */
int omx_randsvd_factors(int64_t n, uint64_t seed, double kappa, int nrefl,
                        double *sigma, double *P, double *Q, void *stream);

/* Bytes of scratch omx_randsvd_f32 needs for order n. */
size_t omx_randsvd_workspace_bytes(int64_t n);

/* A (FP32, RNE of the FP64 entries) = diag(sigma) + P Q^T; diag (FP64, n)
 * receives A_ii and rowabs (FP64, n) the row sums of |A_ij| (their max is
 * ||A||_inf). */
int omx_randsvd_f32(int64_t n, int64_t rank, const double *sigma,
                    const double *P, const double *Q, float *A, int64_t lda,
                    double *diag, double *rowabs, void *work,
                    size_t work_bytes, void *stream);

/* ---------------------------------------------------------------------------
 * The Fasi-Higham HPL-AI matrix, the reference's --highammat family
 * (main.cpp:239-240, highammgen.hpp:1-31, higham_mat_impl.cpp:37-160, after
 * github.com/higham/hpl-ai-matrix): A = L U with L unit lower, l_ij = -alpha
 * below the diagonal, and U unit upper, u_ij = -beta above it, so
 *   a_ij = -beta + alpha beta i (i < j),  1 + alpha beta i (i = j),
 *          -alpha + alpha beta j (i > j)          (0-based i, j),
 * computed in FP64 as written (product, then sum).  The no-pivot LU is
 * exact; alpha = rho beta with beta set so that the paper's estimate of
 * cond_inf(A) equals kappa (higham_mat_impl.cpp:143-148).  alpha and beta
 * are signed: the reference's panel generator (matgen.hpp:157-208, with
 * hmg_gemv_*, otf_gemv.cpp:135-316, and higham_infnorm,
 * panel_norm.hpp:137-192) builds L = I + alpha N, U = I + beta N, which is
 * this matrix at (-alpha, -beta).
 * ------------------------------------------------------------------------- */
/* beta for order n, target kappa and rho = alpha / beta (bisection on the
 * paper's cond_inf estimate; higham_mat_comp_beta, higham_mat_impl.cpp:37-128,
 * with its integer products in 64 bits); NaN if no beta attains kappa. */
double omx_higham_beta(int64_t n, double kappa, double rho);
/* The paper's cond_inf(A) estimate for (alpha, beta) (fhpl,
 * higham_mat_impl.cpp:37-66). */
double omx_higham_cond_est(int64_t n, double alpha, double beta);
/* ||A||_inf of the FP64 matrix (host, O(n); higham_infnorm,
 * panel_norm.hpp:137-192). */
double omx_higham_norm_inf(int64_t n, double alpha, double beta);
/* A (device, FP32, column-major, lda >= n) = the matrix rounded to FP32;
 * diag (device, n) = its FP64 diagonal (hplai_matrix_impl,
 * higham_mat_impl.cpp:130-141). */
int omx_higham_matgen_f32(int64_t n, double alpha, double beta, float *A,
                          int64_t lda, double *diag, void *stream);
/* y += s A x with the FP64 matrix regenerated on the fly (hmg_gemv_up /
 * hmg_gemv_low, otf_gemv.cpp:135-316); x, y device. */
int omx_higham_gemv_f64(int64_t n, double alpha, double beta, double s,
                        const double *x, double *y, void *stream);

/* Bytes of scratch omx_lowrank_gemv_f64 needs. */
size_t omx_lowrank_gemv_workspace_bytes(int64_t rank);

/* y += alpha * (diag(sigma) + P Q^T) x in FP64: the residual product of the
 * refinement for a randsvd system (O(n * rank) work). */
int omx_lowrank_gemv_f64(int64_t n, int64_t rank, const double *sigma,
                         const double *P, const double *Q, double alpha,
                         const double *x, double *y, void *work,
                         size_t work_bytes, void *stream);

/* Bytes of scratch the triangular solves need for order n: 0 (they work in
 * place on x); kept so callers written against a workspace-taking solve
 * keep working.  `work` may be NULL. */
size_t omx_trsv_workspace_bytes(int64_t n);

/* x <- L^-1 x with L the unit-lower FP32 factor stored in A (order n), FP64
 * arithmetic.  `b` is the factorisation's block size and only checked
 * (n % b == 0); the solve walks 256-column blocks: per block column the
 * next diagonal block is updated and solved while the rows below receive
 * L_ij x_j, one launch per block column.  Replaces trsvL_h
 * (panel_trsv.hpp:308-449: cast_copy_general + rocblas_dtrsv +
 * HPLMXP_gemv + daxpy) for one GPU, without the reference's per-block-column
 * swap of accumulation buffers. */
int omx_trsv_lower_unit_f32f64(int64_t n, int64_t b, const float *A,
                               int64_t lda, double *x, void *work,
                               size_t work_bytes, void *stream);

/* x <- U^-1 x with U the upper non-unit FP32 factor in A, block columns
 * from the last to the first.  Replaces trsvU_h (panel_trsv.hpp:452-597)
 * for one GPU. */
int omx_trsv_upper_f32f64(int64_t n, int64_t b, const float *A, int64_t lda,
                          double *x, void *work, size_t work_bytes,
                          void *stream);

/* ---------------------------------------------------------------------------
 * Panel packing and the MFMA Schur GEMM.
 *
 * A panel is an R x K operand (K = the block size b).  Packed panels use the
 * tile layout the GEMM streams: rows in strips of 256; each (strip, K-stage)
 * is one contiguous 32 KB chunk (64 columns for FP16, 128 for FP8); inside
 * it the data are ordered as the lane operands of v_mfma_f32_32x32x16_f16 /
 * v_mfma_scale_f32_32x32x64_f8f6f4 (openmxp_amd/csrc/omx_layout.hpp gives
 * the exact offset of element (r, k)).  Rows beyond R are zero.
 * ------------------------------------------------------------------------- */

/* Bytes of a packed R x K panel (fp8 = 0: FP16, 1: FP8 e4m3). */
size_t omx_panel_bytes(int64_t rows, int64_t k, int fp8);

/* Number of power-of-two exponents (one per 256 x 256 tile) of an FP8
 * panel. */
size_t omx_panel_exps_count(int64_t rows, int64_t k);

/* Pack P = fp16(S) with S(r, l) = src[r*rs + l*ls], r < rows, l < k
 * (k % 64 == 0), round to nearest even, optional triangular mask
 * (OMX_MASK_*).  rs = 1, ls = lda packs a column panel A[k+1:, k]; rs = lda,
 * ls = 1 packs a row panel A[k, k+1:] transposed.  chunk_max (optional,
 * ceil(rows/256) * k/64 floats) receives max |S| per (strip, 64 columns).
 * Replaces half_conversion_left, half_conversion_right_trans and the
 * downCast_copy_* / downCast_trans_general launchers
 * (fp16_gpu_kernels.h:26-49, fp16_gpu_kernels.cpp:11-303). */
int omx_pack_f16(int64_t rows, int64_t k, const float *src, int64_t rs,
                 int64_t ls, int mask, uint16_t *P, float *chunk_max,
                 void *stream);

/* Pack P = e4m3(S * 2^e) with one exponent e per 256 x 256 tile, the
 * largest with max|tile| * 2^e <= 448 (k % 256 == 0); exps receives e and
 * tile_max the tile max-abs (omx_panel_exps_count entries each, strip
 * major).  No reference counterpart: the reference's conversion takes an
 * unused `scale` argument (panel_trf.hpp:259, :306). */
int omx_pack_f8(int64_t rows, int64_t k, const float *src, int64_t rs,
                int64_t ls, uint8_t *P, int32_t *exps, float *tile_max,
                void *stream);

/* Pack P = fp16(S * 2^e_s) with one power-of-two exponent per 256-row
 * strip (k % 256 == 0): e_s is the largest with max|strip| * 2^e_s < 2^15,
 * so no value overflows FP16 and small panels stay clear of its subnormal
 * range; sexp receives e_s (omx_strips(rows) ints), tile_max the 256 x 256
 * tile maxima (recomputed unless have_max).  Pow2 scaling is exact: where
 * the plain cast neither overflows nor goes subnormal the GEMM result is the
 * same.  No reference counterpart (its conversions ignore their `scale`
 * argument, panel_trf.hpp:259, :306). */
int omx_pack_f16_scaled(int64_t rows, int64_t k, const float *src,
                        int64_t rs, int64_t ls, int have_max, uint16_t *P,
                        int32_t *sexp, float *tile_max, void *stream);

/* omx_pack_f8 with `headroom` (0..3) bits below the tile maximum: e is the
 * largest with max|tile| * 2^e <= 448 * 2^-headroom.  have_max = 1: tile_max
 * already holds the tile maxima (from omx_panel_tile_maxabs) and is not
 * recomputed.  No reference counterpart. */
int omx_pack_f8_ex(int64_t rows, int64_t k, const float *src, int64_t rs,
                   int64_t ls, int headroom, int have_max, uint8_t *P,
                   int32_t *exps, float *tile_max, void *stream);

/* K windows: a packed panel of ldk columns can hold several k-column
 * panels side by side (columns [k0, k0 + k) of it).  omx_pack_f16_window /
 * omx_pack_f8_window write one panel into such a window (FP16: plain RNE,
 * or with per-strip scales as omx_pack_f16_scaled when sexp is given; FP8:
 * as omx_pack_f8_ex, exps laid out as [strip][ldk/256], tile_max
 * panel-local); the _window GEMMs multiply the windows [x_k0, x_k0 + k) and
 * [y_k0, y_k0 + k).  This lets the LU apply two consecutive panels to the
 * trailing matrix in one product (k = 2b).  k0 and ldk are multiples of 64
 * (FP16) or 256 (FP8). */
int omx_pack_f16_window(int64_t rows, int64_t k, const float *src,
                        int64_t rs, int64_t ls, int have_max, uint16_t *P,
                        int32_t *sexp, float *tile_max, int64_t ldk,
                        int64_t k0, void *stream);
int omx_pack_f8_window(int64_t rows, int64_t k, const float *src, int64_t rs,
                       int64_t ls, int headroom, int have_max, uint8_t *P,
                       int32_t *exps, int64_t ldk, int64_t k0,
                       float *tile_max, void *stream);

/* max |S| of each 256 x 256 tile of the panel S (k % 256 == 0),
 * omx_panel_exps_count(rows, k) floats, strip major.  No reference
 * counterpart (feeds the FP8 scales and the precision scheduler). */
int omx_panel_tile_maxabs(int64_t rows, int64_t k, const float *src,
                          int64_t rs, int64_t ls, float *tile_max,
                          void *stream);

/* C = beta*C + alpha * X(x_row0 : x_row0+m, :) * Y(y_row0 : y_row0+n, :)^T
 * with packed FP16 panels X, Y of k columns (k % 64 == 0), FP32
 * accumulation, C column-major FP32 (ldc >= m).  x_row0 and y_row0 are
 * multiples of 256.  beta = 0 does not read C.  Replaces
 * GPUBLAS_SGEMM_EX -> rocblas_gemm_ex(f16, f16 -> f32, compute f32)
 * (rocm_device_macros.h:267-268) at the Schur update
 * (panel_trf_gdirect.hpp:791-795, :836-840, :868-872, :905-909) and the
 * -alt panel products (:158-162, :178-182). */
int omx_gemm_f16_packed(int64_t m, int64_t n, int64_t k, float alpha,
                        const uint16_t *X, int64_t x_row0, const uint16_t *Y,
                        int64_t y_row0, float beta, float *C, int64_t ldc,
                        void *stream);

/* omx_gemm_f16_packed for panels packed by omx_pack_f16_scaled: output
 * tile (i, j) takes alpha * 2^-(xsexp[strip i] + ysexp[strip j]).  Both
 * exponent arrays NULL: unscaled. */
int omx_gemm_f16_packed_scaled(int64_t m, int64_t n, int64_t k, float alpha,
                               const uint16_t *X, const int32_t *xsexp,
                               int64_t x_row0, const uint16_t *Y,
                               const int32_t *ysexp, int64_t y_row0,
                               float beta, float *C, int64_t ldc,
                               void *stream);

/* omx_gemm_f16_packed (x_row0 = y_row0 = 0) for the -alt panel products,
 * whose inverse factor is triangular: tri = 1, Y(j, l) = 0 for l > j (the
 * column panel times the upper-triangular U^-1, panel_trf_gdirect.hpp:100-104,
 * :158-162); tri = 2, X(i, l) = 0 for l > i (the lower-triangular L^-1 times
 * the row panel, :136-140, :178-182); tri = 0: all of K.  Each 256 x 256
 * output tile stops at the end of its diagonal 256-column k tile: the
 * products it skips are exact zeros, so C is that of omx_gemm_f16_packed
 * with about half the MFMAs.  tile_max (beta = 0 only; NULL: none)
 * receives each 256 x 256 tile's max |C| as omx_panel_tile_maxabs stores it
 * for the panel C is: C's rows are the panel's rows for tri 0 and 1, C's
 * columns for tri 2 (the row panel); the same values, without a pass over
 * C. */
int omx_gemm_f16_packed_tri(int64_t m, int64_t n, int64_t k, float alpha,
                            const uint16_t *X, const uint16_t *Y, int tri,
                            float beta, float *C, int64_t ldc,
                            float *tile_max, void *stream);

/* omx_gemm_f16_packed_scaled on K windows of wider packed panels. */
int omx_gemm_f16_packed_window(int64_t m, int64_t n, int64_t k, float alpha,
                               const uint16_t *X, const int32_t *xsexp,
                               int64_t x_row0, int64_t x_ldk, int64_t x_k0,
                               const uint16_t *Y, const int32_t *ysexp,
                               int64_t y_row0, int64_t y_ldk, int64_t y_k0,
                               float beta, float *C, int64_t ldc,
                               void *stream);

/* omx_gemm_f8_packed on K windows of wider packed panels. */
int omx_gemm_f8_packed_window(int64_t m, int64_t n, int64_t k, float alpha,
                              const uint8_t *X, const int32_t *xexp,
                              int64_t x_row0, int64_t x_ldk, int64_t x_k0,
                              const uint8_t *Y, const int32_t *yexp,
                              int64_t y_row0, int64_t y_ldk, int64_t y_k0,
                              float beta, float *C, int64_t ldc, void *stream);

/* As omx_gemm_f16_packed with FP8 e4m3 panels and their tile exponents
 * (k % 256 == 0); the exponents enter the MFMA as E8M0 block scales, so each
 * product is exact and only the FP32 accumulation rounds.  No reference
 * counterpart (FP8 is new). */
int omx_gemm_f8_packed(int64_t m, int64_t n, int64_t k, float alpha,
                       const uint8_t *X, const int32_t *xexp, int64_t x_row0,
                       const uint8_t *Y, const int32_t *yexp, int64_t y_row0,
                       float beta, float *C, int64_t ldc, void *stream);

/* ---------------------------------------------------------------------------
 * FP32 dense kernels of the panel factorisation.
 * ------------------------------------------------------------------------- */

/* C = beta*C + alpha * A * B, FP32, no transposes (k % 16 == 0). */
int omx_sgemm_f32(int64_t m, int64_t n, int64_t k, float alpha,
                  const float *A, int64_t lda, const float *B, int64_t ldb,
                  float beta, float *C, int64_t ldc, void *stream);

/* In-place no-pivot LU of the n x n block A (n % 64 == 0): unit-lower L
 * below the diagonal, U on and above it.  *info (device, optional) is set to
 * the 1-based index of the first zero pivot, or INT_MAX if none.  Replaces
 * GPUSOLVER_SGETRF -> rocsolver_sgetrf_npvt (rocm_device_macros.h:290-291,
 * panel_trf_gdirect.hpp:709-711, :800-801). */
int omx_getrf_nopiv_f32(int64_t n, float *A, int64_t lda, int *info,
                        void *stream);

/* Triangular solve with multiple right-hand sides (variant OMX_TRSM_*),
 * B m x n; the triangle's order (n for RUNN/RLNU, m for LLNU) is a
 * multiple of 64.  Replaces GPUBLAS_STRSM -> rocblas_strsm
 * (rocm_device_macros.h:248-249) in update_left/right_panels_wGPU
 * (panel_trf.hpp:188-246) and the -alt 1 inverses on the identity
 * (panel_trf_gdirect.hpp:89-92, :122-125). */
int omx_trsm_f32(int variant, int64_t m, int64_t n, const float *T,
                 int64_t ldt, float *B, int64_t ldb, void *stream);

/* X = T^-1 as the triangular solve of the identity, X <- I * T^-1, for
 * variant OMX_TRSM_RUNN (T upper, non-unit) or OMX_TRSM_RLNU (T lower,
 * unit); n is a multiple of 64 and at most 65535.  Only the rows of each
 * 64-column block that can be non-zero are solved (the inverse of a
 * triangle is a triangle): those entries equal omx_identity_f32 followed by
 * omx_trsm_f32 bit for bit, the rest are the identity's zeros.  Replaces
 * gen_identity_mat + GPUBLAS_STRSM on the identity in the -alt 1/2 panel
 * updates (panel_trf_gdirect.hpp:87-92, :120-125). */
int omx_trsm_identity_f32(int variant, int64_t n, const float *T, int64_t ldt,
                          float *X, int64_t ldx, void *stream);

/* Both inverses of a factored no-pivot LU block D = L U (n x n, n a
 * multiple of 64, at most 65535; L unit lower in D's strict lower part, U
 * upper): Xu = U^-1 and Xl = L^-1 (ldx >= n; the other triangle of each is
 * zero, Xl's diagonal 1).  The 64 x 64 diagonal blocks are inverted by
 * substitution, then adjacent ranges are merged level by level with two
 * batched FP32 products each (X12 = -X1 U12 X2, X21 = -X2 L21 X1): about
 * 2 log2(n/64) launches instead of the 2 n/64 dependent steps of a solve
 * against the identity.  work: n x n floats.  Replaces the two GPUBLAS_STRSM
 * solves of the identity in the -alt 1/2 panel updates
 * (panel_trf_gdirect.hpp:87-92, :120-125); the result differs from those
 * solves by FP32 rounding only. */
int omx_trtri_lu_f32(int64_t n, const float *D, int64_t ldd, float *Xu, float *Xl,
                     int64_t ldx, float *work, void *stream);

/* A = I (n x n, n <= 65535).  Replaces gen_identity_mat
 * (fp16_gpu_kernels.cpp:305-337). */
int omx_identity_f32(int64_t n, float *A, int64_t lda, void *stream);

/* B = A for an m x n block (n <= 65535).  Replaces the GPU_MEMCPY_2D of the
 * diagonal block (panel_trf_gdirect.hpp:805-811). */
int omx_copy_block_f32(int64_t m, int64_t n, const float *A, int64_t lda,
                       float *B, int64_t ldb, void *stream);

/* ---------------------------------------------------------------------------
 * Learned precision schedule: an MLP 16-64-64-3 (FP32, ReLU) that picks, per
 * LU step, FP16 or FP8 for that step's Schur panels (logits z0, z1) and the
 * FP8 tile-scale headroom (z2, bits, rounded and clamped to 0..3).  No
 * reference counterpart.  Parameters: one FP32 array of OMX_MLP_NPARAMS,
 *   W1[16][64] | b1[64] | W2[64][64] | b2[64] | W3[64][3] | b3[3]
 * (W[i][j]: input i -> unit j).  Loss per sample: softmax cross-entropy of
 * (z0, z1) against the class y (0 FP16, 1 FP8) + 0.5 m (z2 - t)^2; the
 * batch loss is the mean.
 * ------------------------------------------------------------------------- */
#define OMX_MLP_NPARAMS 5443
#define OMX_MLP_NFEATURES 16

/* H1, H2 (batch x 64) and Z (batch x 3) for the inputs X (batch x 16),
 * all row-major. */
int omx_mlp_forward_f32(int64_t batch, const float *params, const float *X,
                        float *H1, float *H2, float *Z, void *stream);

/* Bytes of scratch omx_mlp_backward_f32 needs. */
size_t omx_mlp_backward_workspace_bytes(int64_t batch);

/* grad (OMX_MLP_NPARAMS) and *loss (device float) of the mean loss, from a
 * forward pass's H1, H2, Z; y (int32), t, m (float) per sample. */
int omx_mlp_backward_f32(int64_t batch, const float *params, const float *X,
                         const float *H1, const float *H2, const float *Z,
                         const int32_t *y, const float *t, const float *m,
                         float *grad, float *loss, void *work,
                         size_t work_bytes, void *stream);

/* One SGD step on params: with momentum mu > 0, v <- mu v + g and
 * p <- p - lr v (vel: OMX_MLP_NPARAMS floats, zero at the start); with
 * mu = 0, p <- p - lr g (vel may be NULL).  Each operation rounded in FP32
 * as written (no fused multiply-add).  The plain optimiser of the
 * scheduler's training; omx_mlp_adam_f32 is the one the committed training
 * uses. */
int omx_mlp_sgd_f32(int64_t nparams, float *params, const float *grad, float *vel,
                    float lr, float momentum, void *stream);

/* One Adam step (bias-corrected, step >= 1) on params with moments mom, vel. */
int omx_mlp_adam_f32(int64_t nparams, float *params, const float *grad,
                     float *mom, float *vel, float lr, float beta1,
                     float beta2, float eps, int step, void *stream);

/* The 16 scheduler inputs of LU step `step` (features, device): log2
 * statistics (max, min, mean, spread) of the column and row panels' tile
 * maxima, log2 max/min |u_ii| of the factored diagonal block Akk (order b),
 * the step's position, the trailing size, the panels' size relative to the
 * pivots, and the previous step's choice `prev`. */
int omx_sched_features_f32(int64_t ntiles_l, const float *lmax,
                           int64_t ntiles_r, const float *rmax,
                           const float *Akk, int64_t lda, int64_t b,
                           int64_t step, int64_t nb, int64_t n, int prev,
                           float *features, void *stream);

/* ---------------------------------------------------------------------------
 * Solver: the blocked LU with look-ahead and the FP64 refinement, driven by
 * the host.  A handle owns two HIP streams, events and the panel and vector
 * workspaces, GMRES's basis included (allocated once by omx_solver_create,
 * or enlarged by omx_solver_reserve_gmres; the factorisation and refinement
 * calls allocate nothing).  Calls on one handle must not overlap.
 * ------------------------------------------------------------------------- */
typedef struct omx_solver omx_solver;

#define OMX_SOLVER_FP8 1     /* flag of omx_solver_create: also allocate FP8 panels */
#define OMX_SOLVER_SCALE16 2 /* flag: FP16 Schur panels with per-strip pow2 scales */
#define OMX_SOLVER_GROUP4 4  /* flag: one trailing update per 4 panels (depth 4b) instead of per 2 */
#define OMX_SOLVER_TIMERS 8  /* flag: phase timers (HIP events around every timed phase) */

/* Create a solver for order n, block b (n % b == 0, b % 256 == 0), panel
 * method alt (0: FP32 TRSM, 1: FP16 GEMM with the TRSM-on-identity inverse,
 * 2: as 1 with the triangular masks of the reference's -alt 2).  Replaces
 * build_panels_wGPU / build_workspace (panel.hpp:177-187, :361-454). */
int omx_solver_create(int64_t n, int64_t b, int alt, int flags,
                      omx_solver **out);
int omx_solver_destroy(omx_solver *s);

/* GMRES-IR's basis: omx_solver_create sizes it for restarts up to
 * OMX_GMRES_RESTART_DEFAULT ((restart + 1) n doubles: 82 MB at n = 204800).
 * A larger restart (up to OMX_GMRES_RESTART_MAX) needs this call first; it
 * allocates (set-up, not a hot call) and omx_ir_solve then returns
 * OMX_ERR_WORKSPACE for restarts above the reserved one.  The reference has
 * no GMRES (its refinement is classical, iterative_refinement.hpp:141-171). */
#define OMX_GMRES_RESTART_DEFAULT 50
#define OMX_GMRES_RESTART_MAX 512
int omx_solver_reserve_gmres(omx_solver *s, int restart);

/* In-place no-pivot LU of the order-n matrix A (device, lda >= n,
 * lda % 4 == 0) in the step order of panel_lu_async_wGPU_gdirect_alt2
 * (panel_trf_gdirect.hpp:593-1192) on one GPU, with the diagonal/panel work
 * of step k overlapped with the trailing update of step k on a second
 * stream.  sched (host, n/b ints, optional) fixes each step's Schur panels:
 * 0 FP16, 1 + h scaled FP8 e4m3 with tile-scale headroom h (0..3; FP8
 * requires OMX_SOLVER_FP8), -1 FP32 (the step's panels by FP32 TRSM and its
 * Schur products on the FP32 matrix cores straight from A: the precision
 * class for systems whose condition number times the FP16 unit roundoff is
 * not small; no counterpart in the reference, whose Schur update is always
 * FP16).  sched = NULL: the learned policy if one was set
 * (omx_solver_set_policy), otherwise FP16 throughout.  Blocks until done.
 * timings (host, optional, 5 doubles): wall ms, summed ms of the
 * trailing-update GEMMs (HIP events on their stream), their count, their
 * flops, flops of the panel GEMMs.  Returns 0 or the 1-based index of the
 * first zero pivot. */
int omx_lu_factor(omx_solver *s, float *A, int64_t lda, const int *sched,
                  double *timings);

/* The reference's -skip bit-mask (main.cpp:68-70, applied in
 * panel_trf_gdirect.hpp:707,833,901 and panel_trf.hpp:198,237): bit 0 skips
 * the refinement (the driver's business: the library ignores it), bit 1 the
 * diagonal getrf, bit 2 the panel solves (TRSM or -alt inverse products),
 * bit 3 the Schur GEMMs.  The factors are then wrong: a fault-injection and
 * timing aid, as in the reference.  mask 0 (the default) does everything. */
#define OMX_SKIP_IR 1
#define OMX_SKIP_GETRF 2
#define OMX_SKIP_TRSM 4
#define OMX_SKIP_GEMM 8
int omx_solver_set_skip(omx_solver *s, int mask);

/* After omx_lu_factor: ms[k] (host, n/b doubles) = milliseconds from the
 * start of the factorisation until every Schur product of step k's panels
 * was done (the steps of one group share a time).  The reference prints a
 * projected-rate line from the same quantity every -progress steps, after a
 * device synchronisation per step (panel_trf_gdirect.hpp:1146-1161); here
 * the times come from events and the driver prints the lines afterwards. */
int omx_solver_step_times(omx_solver *s, double *ms);

/* Set (params: host array of OMX_MLP_NPARAMS floats) or clear (NULL) the
 * learned schedule used by omx_lu_factor when sched is NULL.  Needs a solver
 * created with OMX_SOLVER_FP8. */
int omx_solver_set_policy(omx_solver *s, const float *params);

/* The last factorisation's per-step choices (sched, n/b ints: 0 FP16,
 * 1 + h FP8, -1 FP32) and, for FP8 solvers, the 16 scheduler features of every step
 * (features, n/b x 16 floats, row-major; zero for the last step, which has
 * no panels).  Either pointer may be NULL. */
int omx_solver_last_schedule(omx_solver *s, int *sched, float *features);

/* Phase timers of a solver created with OMX_SOLVER_TIMERS: GPU seconds and
 * operation counts per phase of the last omx_lu_factor and the refinement
 * calls since it (omx_lu_factor resets them, as the reference's
 * Timer::initialize at the start of a run), numbered and named as the
 * reference's Timer items (timer.hpp:57-81, names :120-145; replaces
 * Timer::show / dump_mp, :160-220).  One GPU has no broadcasts, so the
 * *_BCAST, *_COMM, TEST and WAIT items stay 0.  Timed: DIAG_LU (the
 * diagonal block's LU and -alt inverses), TRSM_L / TRSM_R (the column /
 * row panel: FP32 TRSM or -alt inverse product), CONV_L / CONV_R (their
 * FP16/FP8 packs), GEMM_UPDATE (every Schur product: trailing, look-ahead
 * and within a group), IR_GEMV, IR_TRSV; TOTAL is the wall time of those
 * calls.  Phases on the two streams overlap, so the items can sum to more
 * than TOTAL.  seconds, ops: host arrays of OMX_TIMER_ITEMS doubles.
 * Returns OMX_ERR_ARG when the solver has no timers. */
#define OMX_TIMER_ITEMS 23
int omx_solver_timers(const omx_solver *s, double *seconds, double *ops);
/* Name of timer item i (0 <= i < OMX_TIMER_ITEMS), or NULL. */
const char *omx_timer_name(int item);

/* FP64 refinement of A x = rhs with the FP32 factors LU (device) and the
 * generated A (regenerated from seed; diag = its FP64 diagonal).
 * method 0: classical IR (iterative_refinement2 GPU branch,
 * iterative_refinement.hpp:141-171, x0 = rhs./diag, stop at harness < 16);
 * method 1: GMRES-IR, the same outer loop with each correction solved by
 * GMRES(restart) on (LU)^-1 A d = (LU)^-1 r to relative tolerance
 * inner_tol.  The loop stops when the harness falls below stop_harness
 * (16 in the reference, iterative_refinement.hpp:163); stop_harness <= 0
 * runs exactly maxit residual evaluations (refinement to the FP64 floor).
 * x_given = 1 starts from x instead of x0 = rhs./diag.
 * Blocks until done.  out (host, 4 doubles): final residual,
 * harness, corrections applied (-1 if not converged), total inner
 * iterations.  trace (host, 2*maxit, optional): residual and harness per
 * step.  Returns 0, OMX_ERR_NOCONV, OMX_ERR_ARG for a method other than 0
 * or 1 (checked before any work is queued), or OMX_ERR_WORKSPACE for a
 * GMRES restart above the reserved capacity. */
int omx_ir_solve(omx_solver *s, const float *LU, int64_t lda, uint64_t seed,
                 const double *diag, const double *rhs, double *x,
                 int method, int maxit, int restart, double inner_tol,
                 double stop_harness, int x_given, double *out,
                 double *trace);

/* The system a refinement solves.  kind OMX_SYSTEM_LCG: the reference
 * generator (seed; ||A||_inf = 2 max diag, panel_norm.hpp:129-135).
 * kind OMX_SYSTEM_LOWRANK: diag(sigma) + P Q^T (rank, sigma, P, Q device;
 * norm_inf = ||A||_inf given by the caller).  kind OMX_SYSTEM_HIGHAM: the
 * Fasi-Higham matrix (alpha, beta; norm_inf given by the caller, e.g.
 * omx_higham_norm_inf).  diag (device, n) is A's FP64 diagonal for every
 * kind (x0 = rhs ./ diag). */
#define OMX_SYSTEM_LCG 0
#define OMX_SYSTEM_LOWRANK 1
#define OMX_SYSTEM_HIGHAM 2
typedef struct omx_system {
    int kind;
    uint64_t seed;
    const double *diag;
    double norm_inf;
    int64_t rank;
    const double *sigma;
    const double *P;
    const double *Q;
    double alpha, beta; /* OMX_SYSTEM_HIGHAM */
} omx_system;

/* omx_ir_solve for any omx_system (same arguments and results). */
int omx_ir_solve_system(omx_solver *s, const float *LU, int64_t lda,
                        const omx_system *sys, const double *rhs, double *x,
                        int method, int maxit, int restart, double inner_tol,
                        double stop_harness, int x_given, double *out,
                        double *trace);

/* hpl-harness ||r||/(||A|| ||x|| + ||b||)/(n * DBL_EPSILON/2) on the host
 * (iterative_refinement.hpp:154-157). */
double omx_harness(double normr, double norma, double normx, double normb,
                   int64_t n);

#ifdef __cplusplus
}
#endif

#endif /* OMXP_H */
