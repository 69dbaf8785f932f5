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

/* Bytes of scratch the blocked triangular solves need for order n. */
size_t omx_trsv_workspace_bytes(int64_t n);

/* x <- L^-1 x with L the unit-lower FP32 factor stored in A (order n, block
 * size b, n % b == 0), FP64 arithmetic: per block column the accumulated
 * contributions are subtracted, the diagonal block is solved, and the
 * blocks below gain L_ij x_j.  Replaces trsvL_h (panel_trsv.hpp:308-449:
 * cast_copy_general + rocblas_dtrsv + HPLMXP_gemv + daxpy) for one GPU,
 * without the reference's per-block-column swap of accumulation buffers. */
int omx_trsv_lower_unit_f32f64(int64_t n, int64_t b, const float *A,
                               int64_t lda, double *x, void *work,
                               size_t work_bytes, void *stream);

/* x <- U^-1 x with U the upper non-unit FP32 factor in A.  Replaces trsvU_h
 * (panel_trsv.hpp:452-597) for one GPU. */
int omx_trsv_upper_f32f64(int64_t n, int64_t b, const float *A, int64_t lda,
                          double *x, void *work, size_t work_bytes,
                          void *stream);

/* hpl-harness ||r||/(||A|| ||x|| + ||b||)/(n * DBL_EPSILON/2) on the host
 * (iterative_refinement.hpp:154-157). */
double omx_harness(double normr, double norma, double normx, double normb,
                   int64_t n);

#ifdef __cplusplus
}
#endif

#endif /* OMXP_H */
