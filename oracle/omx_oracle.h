/*
 * omx_oracle.h -- CPU oracle for the mixed-precision no-pivot LU + iterative
 * refinement path (HPL-MxP style).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this code, and only as the checker or
 * as the timed CPU baseline.  The product library (openmxp_amd/csrc) never
 * links, loads or calls it.
 *
 * Conventions: column-major storage, 64-bit sizes, FP16 values carried as
 * uint16_t bit patterns and FP8 e4m3 values as uint8_t bit patterns.
 *
 * Parity status of each function is listed in omx_oracle.c next to the
 * function; every function there is pinned by at least one independent check
 * in tests/ (compiled reference sources, closed forms, LAPACK, brute force).
 */
#ifndef OMX_ORACLE_H
#define OMX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- threading ---------------------------------------------------------- */
void orc_set_threads(int nthreads);
int orc_get_threads(void);

/* ---- 64-bit LCG (hpl_rand.hpp:12-78) ------------------------------------ */
void orc_lcg_jump(uint64_t steps, uint64_t *a, uint64_t *c);
uint64_t orc_lcg_init(uint64_t seed);
uint64_t orc_lcg_state(uint64_t seed, uint64_t n, uint64_t i, uint64_t j);
double orc_lcg_value(uint64_t state);
float orc_lcg_value_f32(uint64_t state);

/* ---- generators (matgen.hpp:8-117, :419-436; hpl_rand.hpp:200-214) ------- */
void orc_matgen_block(int64_t n, uint64_t seed, int64_t i0, int64_t j0,
                      int64_t m, int64_t k, float *A, int64_t lda);
void orc_diag_gpu_order(int64_t n, uint64_t seed, int64_t i0, int64_t m,
                        double *d);
void orc_diag_seq_order(int64_t n, uint64_t seed, int64_t i0, int64_t m,
                        double *d);
void orc_rhs(int64_t n, uint64_t seed, int64_t i0, int64_t m, double *b);
void orc_matgen_full(int64_t n, uint64_t seed, float *A, int64_t lda,
                     double *diag);

/* ---- precision conversions (fp16_gpu_kernels.cpp:11-94) ------------------ */
uint16_t orc_f32_to_f16(float f);
float orc_f16_to_f32(uint16_t h);
uint8_t orc_f32_to_e4m3(float f);
float orc_e4m3_to_f32(uint8_t q);
void orc_f32_to_f16_array(int64_t n, const float *x, uint16_t *y);
void orc_f16_to_f32_array(int64_t n, const uint16_t *x, float *y);
void orc_f32_to_e4m3_array(int64_t n, const float *x, uint8_t *y);
void orc_e4m3_to_f32_array(int64_t n, const uint8_t *x, float *y);
void orc_cast_f16(int64_t m, int64_t k, const float *A, int64_t lda,
                  uint16_t *H, int64_t ldh);
int orc_pow2_scale_exp(float maxabs, float fmax_target);

/* ---- dense kernels ------------------------------------------------------- */
int orc_getrf_nopiv(int64_t n, float *A, int64_t lda);
void orc_trsm_runn(int64_t m, int64_t n, const float *U, int64_t ldu, float *B,
                   int64_t ldb);
void orc_trsm_llnu(int64_t m, int64_t n, const float *L, int64_t ldl, float *B,
                   int64_t ldb);
void orc_trsm_rlnu(int64_t m, int64_t n, const float *L, int64_t ldl, float *B,
                   int64_t ldb);
int orc_trtri_upper_nonunit(int64_t n, float *A, int64_t lda);
void orc_trtri_lower_unit(int64_t n, float *A, int64_t lda);
void orc_gemm_f16(int64_t m, int64_t n, int64_t k, float alpha,
                  const uint16_t *X, int64_t ldx, const uint16_t *Y,
                  int64_t ldy, float beta, float *C, int64_t ldc);
void orc_gemm_f8(int64_t m, int64_t n, int64_t k, float alpha,
                 const uint8_t *X, int64_t ldx, const int32_t *xexp,
                 const uint8_t *Y, int64_t ldy, const int32_t *yexp,
                 int64_t tile, float beta, float *C, int64_t ldc);

/* ---- LU driver (panel_trf_gdirect.hpp:593-1192, 1x1 grid) ---------------- */
/* sched: per panel step k (0..nb-1), 0 = FP16 Schur, 1 = FP8 e4m3 scaled.
 * May be NULL (all FP16).  Returns 0, or 1+k when step k's pivot block has a
 * zero pivot.  stats (optional, nb x 4 doubles): per step max|L|, max|R|,
 * max|A_kk| after update, min|u_ii|. */
int orc_lu_factor(int64_t n, int64_t b, int alt, float *A, int64_t lda,
                  const int *sched, int64_t fp8_tile, double *stats);

/* ---- refinement (iterative_refinement.hpp:98-174 and helpers) ----------- */
void orc_otf_gemv(int64_t n, uint64_t seed, const double *diag, double alpha,
                  const double *x, double *y);
double orc_norm_inf(int64_t n, const double *x);
double orc_harness(double normr, double norma, double normx, double normb,
                   int64_t n);
void orc_trsv_lower_unit(int64_t n, int64_t b, const float *A, int64_t lda,
                         double *x);
void orc_trsv_upper(int64_t n, int64_t b, const float *A, int64_t lda,
                    double *x);
/* Returns the number of corrections applied before the harness fell below
 * 16, or -1 if maxit steps did not converge.  trace (optional, 2*maxit
 * doubles): residual and harness per step. */
int orc_ir_classic(int64_t n, int64_t b, uint64_t seed, const float *LU,
                   int64_t lda, const double *diag, const double *rhs,
                   int maxit, double *x, double *resid, double *harness,
                   double *trace);
int orc_gmres_ir(int64_t n, int64_t b, uint64_t seed, const float *LU,
                 int64_t lda, const double *diag, const double *rhs,
                 int maxit, int restart, double inner_tol, double *x,
                 double *resid, double *harness, double *trace,
                 int *inner_total);

/* ---- randsvd test matrices (no reference counterpart) ------------------- */
/* Dense FP64 A = H_0..H_{k-1} diag(sigma) H_{2k-1}..H_k (column-major,
 * lda = n), sigma_i = kappa^(-i/(n-1)), reflector t from the LCG values at
 * linear indices n^2 + n + t*n + i, normalised; k = nrefl.  O(n^2) per
 * reflector: small n only. */
void orc_randsvd_dense(int64_t n, uint64_t seed, double kappa, int nrefl,
                       double *A);

/* ---- MLP 16-64-64-3 precision scheduler (new) --------------------------- */
/* Parameters, one flat FP32 array of ORC_MLP_NPARAMS:
 *   W1[16][64] | b1[64] | W2[64][64] | b2[64] | W3[64][3] | b3[3]
 * (W[i][j]: input i -> unit j).  Forward: h1 = relu(x W1 + b1),
 * h2 = relu(h1 W2 + b2), z = h2 W3 + b3.  Loss per sample: cross-entropy of
 * softmax(z0, z1) against the class y (0: FP16, 1: FP8) plus
 * 0.5 * m * (z2 - t)^2 (scale head, mask m); the batch loss is the mean. */
#define ORC_MLP_IN 16
#define ORC_MLP_H 64
#define ORC_MLP_OUT 3
#define ORC_MLP_NPARAMS (16 * 64 + 64 + 64 * 64 + 64 + 64 * 3 + 3)
void orc_mlp_forward(int64_t batch, const float *params, const float *X,
                     float *H1, float *H2, float *Z);
/* grad (ORC_MLP_NPARAMS) = d(mean loss)/d(params); returns the mean loss. */
float orc_mlp_backward(int64_t batch, const float *params, const float *X,
                       const int32_t *y, const float *t, const float *m,
                       float *grad);

#ifdef __cplusplus
}
#endif

#endif
