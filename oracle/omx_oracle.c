/*
 * omx_oracle.c -- CPU oracle for the mixed-precision no-pivot LU + iterative
 * refinement path (HPL-MxP style).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this code, and only as the checker or
 * as the timed CPU baseline.  The product library never links, loads or calls
 * it; the HIP path fails loudly when its own library is missing.
 *
 * Every function restates one piece of the reference (at-aaims/OpenMxP) and
 * cites the file:line it follows.  The column "pinned by" says which
 * independent check in tests/ fixes it:
 *
 *   function                  follows                               pinned by
 *   ------------------------  ------------------------------------  ---------------------------------------------
 *   orc_lcg_*                 hpl_rand.hpp:12-78                     compiled hpl_rand.hpp (oracle/_ref), Python
 *                                                                    big-integer restatement, KAT-0 on the GPU
 *   orc_matgen_block          hpl_rand.hpp:80-129, matgen.hpp:67-82  compiled panel_fill_one_with_rand (_ref)
 *   orc_diag_seq_order        hpl_rand.hpp:200-214                   compiled calc_diag (_ref), bit-exact
 *   orc_diag_gpu_order        matgen.hpp:8-65 (512 strided partials, exact-sum bound vs math.fsum and
 *                             tree 512->4, ((s0+s1)+s2)+s3)          compiled calc_diag within n*eps
 *   orc_rhs                   matgen.hpp:419-436                     compiled panel_fill_one_with_rand (_ref)
 *   orc_f32_to_f16/f16_to_f32 fp16_gpu_kernels.cpp:11-36 (RNE)       compiled fp16sim.hpp (_cvtss_sh, _ref),
 *                                                                    numpy float16, exhaustive over f16 values
 *   orc_f32_to_e4m3/e4m3_to_  new (FP8 e4m3fn, RNE, saturating)      torch.float8_e4m3fn conversion
 *   orc_getrf_nopiv           sgetrf_nopiv.cpp:27-61                 compiled sgetrf_nopiv_naive (_ref) bit-exact,
 *                                                                    compiled sgetrf_nopiv_tuned within 1e-4
 *                                                                    (other blocking), LAPACK getrf on
 *                                                                    column-dominant matrices, zero pivot
 *   orc_trsm_*                getrf_nopiv.hpp:7-24 (BLAS strsm)      scipy.linalg.solve_triangular
 *   orc_trtri_*               rocsolver_strtri call sites             numpy.linalg.inv, T*inv(T) = I
 *                             (panel_trf_gdirect.hpp:813-829),
 *                             LAPACK strti2 order
 *   orc_gemm_f16              fp16sim.hpp:222-252 (shgemm)           compiled shgemm (_ref) bit-exact, FP64
 *                                                                    GEMM bound
 *   orc_gemm_f8               new (FP8 Schur with pow2 tile scales)  FP64 GEMM of decoded operands (exact
 *                                                                    products) within accumulation bound
 *   orc_lu_factor             panel_trf_gdirect.hpp:593-1192,        closed form: L*U reproduces A to FP16
 *                             :46-183 (alt 0/1/2)                    tolerance; scipy dgesv forward error
 *   orc_otf_gemv              panel_gemv.hpp:13-179                  dense FP64 product of the generated matrix;
 *                                                                    KAT-0 (crusher_example_32x32.out:186)
 *   orc_trsv_*                panel_trsv.hpp:308-597 (Q=1 fixed)     scipy.linalg.solve_triangular (FP64)
 *   orc_ir_classic            iterative_refinement.hpp:98-174        harness closed form; scipy dgesv
 *   orc_gmres_ir              new (GMRES-IR, restart m)              scipy dgesv forward error
 */
#include "omx_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* threading                                                                 */
/* ------------------------------------------------------------------------ */
void orc_set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* 64-bit LCG, hpl_rand.hpp:12-78                                            */
/* ------------------------------------------------------------------------ */
#define ORC_LCG_A 6364136223846793005ULL /* hpl_rand.hpp:16 */
#define ORC_LCG_C 1ULL

typedef struct {
    uint64_t a, c;
} orc_map_t;

/* composition x -> f(g(x)) of two affine maps (hpl_rand.hpp:18-20) */
static inline orc_map_t map_compose(orc_map_t f, orc_map_t g) {
    orc_map_t r = {f.a * g.a, f.a * g.c + f.c};
    return r;
}

/* the map of `steps` LCG steps, by binary powering (hpl_rand.hpp:31-48) */
static orc_map_t map_pow(uint64_t steps) {
    orc_map_t base = {ORC_LCG_A, ORC_LCG_C};
    orc_map_t result = {1ULL, 0ULL};
    while (steps) {
        if (steps & 1ULL) result = map_compose(result, base);
        base = map_compose(base, base);
        steps >>= 1;
    }
    return result;
}

static inline uint64_t map_apply(orc_map_t f, uint64_t x) { return f.a * x + f.c; }

void orc_lcg_jump(uint64_t steps, uint64_t *a, uint64_t *c) {
    orc_map_t m = map_pow(steps);
    *a = m.a;
    *c = m.c;
}

/* RandStat::initialize (hpl_rand.hpp:54-57) */
uint64_t orc_lcg_init(uint64_t seed) { return ORC_LCG_A * seed + ORC_LCG_C; }

/* state of element (i, j) of the order-n matrix: linear index i + n*j
 * (Matgen::jump, hpl_rand.hpp:249-254).  Column n is the right-hand side. */
uint64_t orc_lcg_state(uint64_t seed, uint64_t n, uint64_t i, uint64_t j) {
    return map_apply(map_pow(i + n * j), orc_lcg_init(seed));
}

/* RandStat::operator double (hpl_rand.hpp:64-65): int64 reinterpretation
 * times (2^-64 - 2^-117), value in [-0.5, 0.5). */
double orc_lcg_value(uint64_t state) {
    return (double)(int64_t)state * 0x1.fffffffffffffp-65;
}

/* RandStat::operator float (hpl_rand.hpp:74-77) */
float orc_lcg_value_f32(uint64_t state) { return (float)orc_lcg_value(state); }

/* ------------------------------------------------------------------------ */
/* generators                                                                */
/* ------------------------------------------------------------------------ */

/* raw LCG block A(i0:i0+m, j0:j0+k) in FP32, diagonal NOT substituted
 * (panel_fill_one_with_rand without calc_diag, hpl_rand.hpp:80-110;
 * fill_panel_dev, matgen.hpp:67-82) */
void orc_matgen_block(int64_t n, uint64_t seed, int64_t i0, int64_t j0,
                      int64_t m, int64_t k, float *A, int64_t lda) {
    const uint64_t x00 = orc_lcg_init(seed);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < k; ++j) {
        uint64_t s = map_apply(map_pow((uint64_t)(i0) + (uint64_t)n * (uint64_t)(j0 + j)), x00);
        float *col = A + j * lda;
        for (int64_t i = 0; i < m; ++i) {
            col[i] = (float)orc_lcg_value(s);
            s = ORC_LCG_A * s + ORC_LCG_C;
        }
    }
}

/* Diagonal d_i = sum_{j != i} |A_ij| in the order of the reference's
 * on-device generator fill_panel_diag_dev (matgen.hpp:26-62): 512 threads,
 * thread t sums columns t, t+512, ... in order; tree 512 -> 256 -> ... -> 4;
 * d = ((s0 + s1) + s2) + s3. */
#define ORC_DIAG_THREADS 512
void orc_diag_gpu_order(int64_t n, uint64_t seed, int64_t i0, int64_t m,
                        double *d) {
    const uint64_t x00 = orc_lcg_init(seed);
    orc_map_t colstep[ORC_DIAG_THREADS];
    for (int t = 0; t < ORC_DIAG_THREADS; ++t) colstep[t] = map_pow((uint64_t)n * (uint64_t)t);
    const orc_map_t jump_thread = map_pow((uint64_t)ORC_DIAG_THREADS * (uint64_t)n);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t r = 0; r < m; ++r) {
        const int64_t g = i0 + r;
        const uint64_t xg = map_apply(map_pow((uint64_t)g), x00);
        double sh[ORC_DIAG_THREADS];
        for (int t = 0; t < ORC_DIAG_THREADS; ++t) {
            uint64_t s = map_apply(colstep[t], xg);
            double acc = 0.0;
            for (int64_t j = t; j < n; j += ORC_DIAG_THREADS) {
                if (j != g) acc += fabs(orc_lcg_value(s));
                s = map_apply(jump_thread, s);
            }
            sh[t] = acc;
        }
        for (int s = ORC_DIAG_THREADS / 2; s >= 4; s >>= 1)
            for (int t = 0; t < s; ++t) sh[t] += sh[t + s];
        d[r] = sh[0] + sh[1] + sh[2] + sh[3];
    }
}

/* calc_diag order (hpl_rand.hpp:200-214): sequential over columns. */
void orc_diag_seq_order(int64_t n, uint64_t seed, int64_t i0, int64_t m,
                        double *d) {
    const uint64_t x00 = orc_lcg_init(seed);
    const orc_map_t jump_n = map_pow((uint64_t)n);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t r = 0; r < m; ++r) {
        const int64_t g = i0 + r;
        uint64_t s = map_apply(map_pow((uint64_t)g), x00);
        double sum = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            if (j != g) sum += fabs(orc_lcg_value(s));
            s = map_apply(jump_n, s);
        }
        d[r] = sum;
    }
}

/* right-hand side b_i = value at linear index n*n + i (panel_colvgen,
 * matgen.hpp:419-436 via panel_fill_one_with_rand with j0 = n) */
void orc_rhs(int64_t n, uint64_t seed, int64_t i0, int64_t m, double *b) {
    uint64_t s = orc_lcg_state(seed, (uint64_t)n, (uint64_t)i0, (uint64_t)n);
    for (int64_t i = 0; i < m; ++i) {
        b[i] = orc_lcg_value(s);
        s = ORC_LCG_A * s + ORC_LCG_C;
    }
}

/* the reference's default (-init 1) system: FP32 matrix from the LCG, FP64
 * diagonal in the on-device order, A_ii = float(d_i)
 * (panel_matgen_dev, matgen.hpp:84-117; fill_panel_diag_dev :60-63) */
void orc_matgen_full(int64_t n, uint64_t seed, float *A, int64_t lda,
                     double *diag) {
    orc_matgen_block(n, seed, 0, 0, n, n, A, lda);
    orc_diag_gpu_order(n, seed, 0, n, diag);
    for (int64_t i = 0; i < n; ++i) A[i + i * lda] = (float)diag[i];
}

/* ------------------------------------------------------------------------ */
/* precision conversions                                                     */
/* ------------------------------------------------------------------------ */
static inline uint32_t f2u(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
static inline float u2f(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

/* round-to-nearest-even right shift of a 32-bit significand */
static inline uint32_t rne_shift(uint32_t m, int s) {
    if (s <= 0) return m << (-s);
    if (s >= 32) return 0;
    uint32_t q = m >> s;
    uint32_t rem = m & ((1u << s) - 1u);
    uint32_t half = 1u << (s - 1);
    if (rem > half || (rem == half && (q & 1u))) ++q;
    return q;
}

/* IEEE binary32 -> binary16, round to nearest even (__float2half in
 * fp16_gpu_kernels.cpp:35; _cvtss_sh(x, 0) in fp16sim.hpp:119) */
uint16_t orc_f32_to_f16(float f) {
    const uint32_t x = f2u(f);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) {
        if (ax > 0x7f800000u) return (uint16_t)(sign | 0x7e00u | ((ax >> 13) & 0x3ffu));
        return (uint16_t)(sign | 0x7c00u);
    }
    const int e = (int)(ax >> 23);
    const uint32_t mant = (ax & 0x7fffffu) | (e ? 0x800000u : 0u);
    if (e >= 113) { /* |f| >= 2^-14: normal binary16 (or overflow) */
        uint32_t h = ((uint32_t)(e - 112) << 10) + ((ax >> 13) & 0x3ffu);
        const uint32_t rem = ax & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
        if (h >= 0x7c00u) h = 0x7c00u;
        return (uint16_t)(sign | h);
    }
    /* subnormal binary16: units of 2^-24; value = mant * 2^(e-150) */
    const int s = (e ? 126 - e : 125);
    return (uint16_t)(sign | rne_shift(mant, s));
}

float orc_f16_to_f32(uint16_t h) {
    const uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu;
    const uint32_t m = h & 0x3ffu;
    if (e == 0x1f) return u2f(sign | 0x7f800000u | (m << 13));
    if (e == 0) {
        float v = (float)m * 0x1p-24f;
        return sign ? -v : v;
    }
    return u2f(sign | ((e + 112u) << 23) | (m << 13));
}

/* binary32 -> OCP FP8 e4m3fn (bias 7, max 448, no infinities), round to
 * nearest even, saturating to +-448; NaN -> 0x7f. */
uint8_t orc_f32_to_e4m3(float f) {
    const uint32_t x = f2u(f);
    const uint8_t sign = (uint8_t)((x >> 24) & 0x80u);
    const uint32_t ax = x & 0x7fffffffu;
    if (ax > 0x7f800000u) return (uint8_t)(sign | 0x7fu);
    if (ax >= 0x43e00000u) return (uint8_t)(sign | 0x7eu); /* >= 448 (and inf) */
    const int e = (int)(ax >> 23);
    const uint32_t mant = (ax & 0x7fffffu) | (e ? 0x800000u : 0u);
    if (e >= 121) { /* |f| >= 2^-6: normal e4m3 */
        uint32_t q = ((uint32_t)(e - 120) << 3) + ((ax >> 20) & 0x7u);
        const uint32_t rem = ax & 0xfffffu;
        if (rem > 0x80000u || (rem == 0x80000u && (q & 1u))) ++q;
        if (q > 0x7eu) q = 0x7eu;
        return (uint8_t)(sign | q);
    }
    /* subnormal e4m3: units of 2^-9; value = mant * 2^(e-150) */
    const int s = (e ? 141 - e : 140);
    return (uint8_t)(sign | rne_shift(mant, s));
}

float orc_e4m3_to_f32(uint8_t q) {
    const int neg = q & 0x80;
    const uint32_t e = (q >> 3) & 0xfu;
    const uint32_t m = q & 0x7u;
    float v;
    if (e == 0xf && m == 0x7) return neg ? -NAN : NAN;
    if (e == 0)
        v = (float)m * 0x1p-9f;
    else
        v = u2f(((e + 120u) << 23) | (m << 20));
    return neg ? -v : v;
}

/* element-wise array forms of the four conversions above */
void orc_f32_to_f16_array(int64_t n, const float *x, uint16_t *y) {
    for (int64_t i = 0; i < n; ++i) y[i] = orc_f32_to_f16(x[i]);
}
void orc_f16_to_f32_array(int64_t n, const uint16_t *x, float *y) {
    for (int64_t i = 0; i < n; ++i) y[i] = orc_f16_to_f32(x[i]);
}
void orc_f32_to_e4m3_array(int64_t n, const float *x, uint8_t *y) {
    for (int64_t i = 0; i < n; ++i) y[i] = orc_f32_to_e4m3(x[i]);
}
void orc_e4m3_to_f32_array(int64_t n, const uint8_t *x, float *y) {
    for (int64_t i = 0; i < n; ++i) y[i] = orc_e4m3_to_f32(x[i]);
}

/* FP32 -> FP16 column-major copy (downCast_copy_general,
 * fp16_gpu_kernels.cpp:215-254) */
void orc_cast_f16(int64_t m, int64_t k, const float *A, int64_t lda,
                  uint16_t *H, int64_t ldh) {
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < k; ++j)
        for (int64_t i = 0; i < m; ++i) H[i + j * ldh] = orc_f32_to_f16(A[i + j * lda]);
}

/* largest e with maxabs * 2^e <= fmax_target (power-of-two tile scale);
 * 0 for an all-zero tile; clamped to [-100, 100]. */
int orc_pow2_scale_exp(float maxabs, float fmax_target) {
    if (!(maxabs > 0.0f) || !isfinite(maxabs)) return 0;
    int em, et;
    const float fm = frexpf(maxabs, &em); /* maxabs = fm * 2^em, fm in [0.5,1) */
    const float ft = frexpf(fmax_target, &et);
    int e = (fm <= ft) ? (et - em) : (et - em - 1);
    if (e > 100) e = 100;
    if (e < -100) e = -100;
    return e;
}

/* ------------------------------------------------------------------------ */
/* dense FP32 kernels                                                        */
/* ------------------------------------------------------------------------ */

/* no-pivot LU, right-looking, multiply by the reciprocal pivot
 * (sgetrf_nopiv_naive, sgetrf_nopiv.cpp:27-61).  Returns 0 or 1 + index of
 * the first zero pivot (LAPACK info convention); computation continues. */
int orc_getrf_nopiv(int64_t n, float *A, int64_t lda) {
    int info = 0;
    for (int64_t k = 0; k < n; ++k) {
        const float akk = A[k + k * lda];
        if (akk == 0.0f && info == 0) info = (int)(k + 1);
        const float inv = 1.0f / akk;
        float *ak = A + k * lda;
        for (int64_t i = k + 1; i < n; ++i) ak[i] *= inv;
#pragma omp parallel for schedule(static) if (n - k > 256)
        for (int64_t j = k + 1; j < n; ++j) {
            float *aj = A + j * lda;
            const float akj = aj[k];
            for (int64_t i = k + 1; i < n; ++i) aj[i] -= ak[i] * akj;
        }
    }
    return info;
}

/* B <- B * U^-1 (Right, Upper, NoTrans, NonUnit); reference BLAS order:
 * column j minus earlier columns, then times 1/U_jj (getrf_nopiv.hpp:11-14) */
void orc_trsm_runn(int64_t m, int64_t n, const float *U, int64_t ldu, float *B,
                   int64_t ldb) {
    const int64_t RB = 256;
#pragma omp parallel for schedule(static)
    for (int64_t i0 = 0; i0 < m; i0 += RB) {
        const int64_t i1 = (i0 + RB < m) ? i0 + RB : m;
        for (int64_t j = 0; j < n; ++j) {
            float *bj = B + j * ldb;
            for (int64_t l = 0; l < j; ++l) {
                const float ulj = U[l + j * ldu];
                if (ulj != 0.0f) {
                    const float *bl = B + l * ldb;
                    for (int64_t i = i0; i < i1; ++i) bj[i] -= ulj * bl[i];
                }
            }
            const float t = 1.0f / U[j + j * ldu];
            for (int64_t i = i0; i < i1; ++i) bj[i] *= t;
        }
    }
}

/* B <- L^-1 * B (Left, Lower, NoTrans, Unit) (getrf_nopiv.hpp:19-22) */
void orc_trsm_llnu(int64_t m, int64_t n, const float *L, int64_t ldl, float *B,
                   int64_t ldb) {
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j) {
        float *bj = B + j * ldb;
        for (int64_t l = 0; l < m; ++l) {
            const float blj = bj[l];
            if (blj != 0.0f) {
                const float *ll = L + l * ldl;
                for (int64_t i = l + 1; i < m; ++i) bj[i] -= blj * ll[i];
            }
        }
    }
}

/* B <- B * L^-1 (Right, Lower, NoTrans, Unit): the -alt 1 inverse of the
 * unit-lower factor, solved on the identity (panel_trf_gdirect.hpp:121-125) */
void orc_trsm_rlnu(int64_t m, int64_t n, const float *L, int64_t ldl, float *B,
                   int64_t ldb) {
    /* element (i, j): b_ij -= L_lj b_il for l = j+1, ..., n-1 in that order,
     * columns j from the last; done on 64-row column slices (contiguous in
     * column-major storage), each element's terms in the same order */
    const int64_t RB = 64;
#pragma omp parallel for schedule(static)
    for (int64_t i0 = 0; i0 < m; i0 += RB) {
        const int64_t i1 = (i0 + RB < m) ? i0 + RB : m;
        for (int64_t j = n - 1; j >= 0; --j) {
            float *bj = B + j * ldb;
            for (int64_t l = j + 1; l < n; ++l) {
                const float llj = L[l + j * ldl];
                const float *bl = B + l * ldb;
                for (int64_t i = i0; i < i1; ++i) bj[i] -= llj * bl[i];
            }
        }
    }
}

/* in-place inverse of the upper non-unit triangle, LAPACK strti2 order
 * (the reference calls rocsolver_strtri, panel_trf_gdirect.hpp:814-820) */
int orc_trtri_upper_nonunit(int64_t n, float *A, int64_t lda) {
    int info = 0;
    for (int64_t j = 0; j < n; ++j) {
        if (A[j + j * lda] == 0.0f && info == 0) info = (int)(j + 1);
        A[j + j * lda] = 1.0f / A[j + j * lda];
        const float ajj = -A[j + j * lda];
        float *x = A + j * lda; /* x = A(0:j, j) */
        /* x := triu(A(0:j,0:j)) * x  (STRMV upper, no-trans, non-unit) */
        for (int64_t j2 = 0; j2 < j; ++j2) {
            const float t = x[j2];
            if (t != 0.0f) {
                const float *c = A + j2 * lda;
                for (int64_t i = 0; i < j2; ++i) x[i] += t * c[i];
                x[j2] = t * c[j2];
            }
        }
        for (int64_t i = 0; i < j; ++i) x[i] *= ajj;
    }
    return info;
}

/* in-place inverse of the strictly-lower part with unit diagonal, LAPACK
 * strti2 order (rocsolver_strtri lower/unit, panel_trf_gdirect.hpp:822-828) */
void orc_trtri_lower_unit(int64_t n, float *A, int64_t lda) {
    for (int64_t j = n - 1; j >= 0; --j) {
        if (j < n - 1) {
            float *x = A + (j + 1) + j * lda; /* x = A(j+1:n, j), length n-j-1 */
            const int64_t len = n - j - 1;
            const float *T = A + (j + 1) + (j + 1) * lda;
            /* x := tril_unit(T) * x  (STRMV lower, no-trans, unit) */
            for (int64_t j2 = len - 1; j2 >= 0; --j2) {
                const float t = x[j2];
                if (t != 0.0f) {
                    const float *c = T + j2 * lda;
                    for (int64_t i = len - 1; i > j2; --i) x[i] += t * c[i];
                }
            }
            for (int64_t i = 0; i < len; ++i) x[i] = -x[i];
        }
    }
}

/* C = beta*C + alpha * X * Y^T with X (m x k), Yt (k x n, column j holds
 * Y(j, 0:k)) given as exact FP32 values; per element the k products are
 * summed in order in FP32 (shgemm, fp16sim.hpp:222-252), starting from 0,
 * then combined with C.  The products of FP16 (or scaled FP8) values are
 * exact in FP32, so fused and unfused multiply-add give the same sums.
 *
 * Full 16 x 6 tiles keep their sums in registers (X repacked into 16-row
 * panels, contiguous in l); edge tiles take the plain loop.  Both add the
 * same terms in the same order, so the tiling does not change any bit. */
enum { GM_MR = 16, GM_NR = 6 };

static void gemm_tile_store(int64_t mi, int64_t nj, const float *T, int64_t ldt, float alpha, float beta,
                            float *C, int64_t ldc) {
    for (int64_t c = 0; c < nj; ++c) {
        float *cc = C + c * ldc;
        const float *t = T + c * ldt;
        if (beta == 0.0f)
            for (int64_t i = 0; i < mi; ++i) cc[i] = t[i] * alpha;
        else
            for (int64_t i = 0; i < mi; ++i) cc[i] = cc[i] * beta + t[i] * alpha;
    }
}

static void gemm_seq_core(int64_t m, int64_t n, int64_t k, float alpha,
                          const float *Xf, const float *Yt, float beta,
                          float *C, int64_t ldc) {
    const int64_t nib = (m + GM_MR - 1) / GM_MR, njb = (n + GM_NR - 1) / GM_NR;
    /* Xp: 16-row panels, panel ib holds X(16 ib + r, l) at [(ib k + l) 16 + r]
     * (rows past m are zero and never stored) */
    float *Xp = (float *)malloc(sizeof(float) * (size_t)(nib * k * GM_MR + 1));
#pragma omp parallel for schedule(static)
    for (int64_t ib = 0; ib < nib; ++ib)
        for (int64_t l = 0; l < k; ++l)
            for (int64_t r = 0; r < GM_MR; ++r) {
                const int64_t i = ib * GM_MR + r;
                Xp[(ib * k + l) * GM_MR + r] = i < m ? Xf[i + l * m] : 0.0f;
            }
    /* row panels outermost: a thread's consecutive tiles share the panel,
     * which stays in cache across the column blocks */
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t ib = 0; ib < nib; ++ib) {
        for (int64_t jb = 0; jb < njb; ++jb) {
            const int64_t i0 = ib * GM_MR, j0 = jb * GM_NR;
            const int64_t mi = (i0 + GM_MR <= m) ? GM_MR : m - i0;
            const int64_t nj = (j0 + GM_NR <= n) ? GM_NR : n - j0;
            const float *xp = Xp + ib * k * GM_MR;
            float T[GM_NR][GM_MR];
            if (nj == GM_NR) {
                float acc[GM_NR][GM_MR];
                memset(acc, 0, sizeof(acc));
                const float *y0 = Yt + j0 * k;
                for (int64_t l = 0; l < k; ++l) {
                    const float *xl = xp + l * GM_MR;
#pragma GCC unroll 6
                    for (int c = 0; c < GM_NR; ++c) {
                        const float y = y0[l + c * k];
#pragma GCC unroll 16
                        for (int r = 0; r < GM_MR; ++r) acc[c][r] = __builtin_fmaf(xl[r], y, acc[c][r]);
                    }
                }
                memcpy(T, acc, sizeof(T));
            } else {
                memset(T, 0, sizeof(T));
                for (int64_t l = 0; l < k; ++l) {
                    const float *xl = xp + l * GM_MR;
                    for (int64_t c = 0; c < nj; ++c) {
                        const float y = Yt[l + (j0 + c) * k];
                        for (int r = 0; r < GM_MR; ++r) T[c][r] = __builtin_fmaf(xl[r], y, T[c][r]);
                    }
                }
            }
            gemm_tile_store(mi, nj, &T[0][0], GM_MR, alpha, beta, C + i0 + j0 * ldc, ldc);
        }
    }
    free(Xp);
}

/* C = beta*C + alpha * X * Y^T, X(i,l) = X[i + l*ldx] (m x k FP16),
 * Y(j,l) = Y[j + l*ldy] (n x k FP16).  The reference's Schur update is this
 * with alpha = -1, beta = 1 (panel_trf_gdirect.hpp:905-909). */
void orc_gemm_f16(int64_t m, int64_t n, int64_t k, float alpha,
                  const uint16_t *X, int64_t ldx, const uint16_t *Y,
                  int64_t ldy, float beta, float *C, int64_t ldc) {
    if (m <= 0 || n <= 0) return;
    float *Xf = (float *)malloc(sizeof(float) * (size_t)(m * k + 1));
    float *Yt = (float *)malloc(sizeof(float) * (size_t)(n * k + 1));
#pragma omp parallel for schedule(static)
    for (int64_t l = 0; l < k; ++l)
        for (int64_t i = 0; i < m; ++i) Xf[i + l * m] = orc_f16_to_f32(X[i + l * ldx]);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j)
        for (int64_t l = 0; l < k; ++l) Yt[l + j * k] = orc_f16_to_f32(Y[j + l * ldy]);
    gemm_seq_core(m, n, k, alpha, Xf, Yt, beta, C, ldc);
    free(Xf);
    free(Yt);
}

/* FP8 variant: X(i,l) = e4m3(X[i + l*ldx]) * 2^-xexp[(i/tile) + (l/tile)*mt]
 * with mt = ceil(m/tile); Y likewise with nt = ceil(n/tile). */
void orc_gemm_f8(int64_t m, int64_t n, int64_t k, float alpha,
                 const uint8_t *X, int64_t ldx, const int32_t *xexp,
                 const uint8_t *Y, int64_t ldy, const int32_t *yexp,
                 int64_t tile, float beta, float *C, int64_t ldc) {
    if (m <= 0 || n <= 0) return;
    const int64_t mt = (m + tile - 1) / tile, nt = (n + tile - 1) / tile;
    float *Xf = (float *)malloc(sizeof(float) * (size_t)(m * k + 1));
    float *Yt = (float *)malloc(sizeof(float) * (size_t)(n * k + 1));
#pragma omp parallel for schedule(static)
    for (int64_t l = 0; l < k; ++l)
        for (int64_t i = 0; i < m; ++i)
            Xf[i + l * m] = ldexpf(orc_e4m3_to_f32(X[i + l * ldx]), -xexp[(i / tile) + (l / tile) * mt]);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j)
        for (int64_t l = 0; l < k; ++l)
            Yt[l + j * k] = ldexpf(orc_e4m3_to_f32(Y[j + l * ldy]), -yexp[(j / tile) + (l / tile) * nt]);
    gemm_seq_core(m, n, k, alpha, Xf, Yt, beta, C, ldc);
    free(Xf);
    free(Yt);
}

/* ------------------------------------------------------------------------ */
/* LU driver                                                                 */
/* ------------------------------------------------------------------------ */

/* A panel of the Schur update in either precision.  rows x b, column-major
 * with ld = rows.  For FP8 the per-(tile x tile) exponents live in exps. */
typedef struct {
    int64_t rows, b;
    int fp8;  /* 1: FP8 e4m3 with tile exponents; -1: FP32 as factored; 0: FP16 */
    int64_t tile;
    uint16_t *h;
    uint8_t *q;
    int32_t *exps;
    float *f;
} orc_panel_t;

static void panel_alloc(orc_panel_t *p, int64_t rows_max, int64_t b, int64_t tile) {
    p->rows = 0;
    p->b = b;
    p->fp8 = 0;
    p->tile = tile;
    p->h = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)(rows_max * b + 1));
    p->q = (uint8_t *)malloc((size_t)(rows_max * b + 1));
    const int64_t tmax = (rows_max + tile - 1) / tile * ((b + tile - 1) / tile) + 1;
    p->exps = (int32_t *)malloc(sizeof(int32_t) * (size_t)tmax);
    p->f = (float *)malloc(sizeof(float) * (size_t)(rows_max * b + 1));
}

static void panel_free(orc_panel_t *p) {
    free(p->h);
    free(p->q);
    free(p->exps);
    free(p->f);
}

/* Fill a panel from FP32 source S(r, l) = src[r*rs + l*ls], r < rows,
 * l < b.  The left panel reads A[k+1:, k] (rs = 1, ls = lda); the right
 * panel reads A[k, k+1:] transposed (rs = lda, ls = 1), as
 * half_conversion_left / half_conversion_right_trans do
 * (panel_trf.hpp:258-325, fp16_gpu_kernels.cpp:157-211). */
static void panel_fill(orc_panel_t *p, int64_t rows, const float *src, int64_t rs,
                       int64_t ls, int fp8, double *maxabs_out) {
    p->rows = rows;
    p->fp8 = fp8;
    const int64_t b = p->b;
    double mx = 0.0;
#pragma omp parallel for schedule(static) reduction(max : mx)
    for (int64_t l = 0; l < b; ++l)
        for (int64_t r = 0; r < rows; ++r) {
            const double v = fabs((double)src[r * rs + l * ls]);
            if (v > mx) mx = v;
        }
    if (maxabs_out) *maxabs_out = mx;
    if (fp8 < 0) {
#pragma omp parallel for schedule(static)
        for (int64_t l = 0; l < b; ++l)
            for (int64_t r = 0; r < rows; ++r) p->f[r + l * rows] = src[r * rs + l * ls];
        return;
    }
    if (!fp8) {
#pragma omp parallel for schedule(static)
        for (int64_t l = 0; l < b; ++l)
            for (int64_t r = 0; r < rows; ++r) p->h[r + l * rows] = orc_f32_to_f16(src[r * rs + l * ls]);
        return;
    }
    const int64_t tile = p->tile;
    const int64_t rt = (rows + tile - 1) / tile, lt = (b + tile - 1) / tile;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t tl = 0; tl < lt; ++tl)
        for (int64_t tr = 0; tr < rt; ++tr) {
            const int64_t r0 = tr * tile, r1 = (r0 + tile < rows) ? r0 + tile : rows;
            const int64_t l0 = tl * tile, l1 = (l0 + tile < b) ? l0 + tile : b;
            float tmx = 0.0f;
            for (int64_t l = l0; l < l1; ++l)
                for (int64_t r = r0; r < r1; ++r) {
                    const float v = fabsf(src[r * rs + l * ls]);
                    if (v > tmx) tmx = v;
                }
            const int e = orc_pow2_scale_exp(tmx, 448.0f);
            p->exps[tr + tl * rt] = e;
            for (int64_t l = l0; l < l1; ++l)
                for (int64_t r = r0; r < r1; ++r)
                    p->q[r + l * rows] = orc_f32_to_e4m3(ldexpf(src[r * rs + l * ls], e));
        }
}

/* C(mr x nc, ldc) -= Lp(rows lr0.., :) * Rp(rows rc0.., :)^T */
static void panel_schur(const orc_panel_t *Lp, int64_t lr0, const orc_panel_t *Rp,
                        int64_t rc0, int64_t mr, int64_t nc, float *C, int64_t ldc) {
    if (mr <= 0 || nc <= 0) return;
    const int64_t b = Lp->b;
    if (!Lp->fp8) {
        orc_gemm_f16(mr, nc, b, -1.0f, Lp->h + lr0, Lp->rows, Rp->h + rc0, Rp->rows, 1.0f, C, ldc);
        return;
    }
    if (Lp->fp8 < 0) {
        /* FP32 class: the factored panels as they are, the same sequential
         * FP32 accumulation */
        float *Xf = (float *)malloc(sizeof(float) * (size_t)(mr * b + 1));
        float *Yt = (float *)malloc(sizeof(float) * (size_t)(nc * b + 1));
#pragma omp parallel for schedule(static)
        for (int64_t l = 0; l < b; ++l)
            for (int64_t i = 0; i < mr; ++i) Xf[i + l * mr] = Lp->f[lr0 + i + l * Lp->rows];
#pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < nc; ++j)
            for (int64_t l = 0; l < b; ++l) Yt[l + j * b] = Rp->f[rc0 + j + l * Rp->rows];
        gemm_seq_core(mr, nc, b, -1.0f, Xf, Yt, 1.0f, C, ldc);
        free(Xf);
        free(Yt);
        return;
    }
    /* FP8: decode both panels with their tile exponents, then the same
     * sequential FP32 accumulation. */
    const int64_t tile = Lp->tile;
    const int64_t lrt = (Lp->rows + tile - 1) / tile, rrt = (Rp->rows + tile - 1) / tile;
    float *Xf = (float *)malloc(sizeof(float) * (size_t)(mr * b + 1));
    float *Yt = (float *)malloc(sizeof(float) * (size_t)(nc * b + 1));
#pragma omp parallel for schedule(static)
    for (int64_t l = 0; l < b; ++l)
        for (int64_t i = 0; i < mr; ++i) {
            const int64_t r = lr0 + i;
            Xf[i + l * mr] = ldexpf(orc_e4m3_to_f32(Lp->q[r + l * Lp->rows]), -Lp->exps[r / tile + (l / tile) * lrt]);
        }
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < nc; ++j)
        for (int64_t l = 0; l < b; ++l) {
            const int64_t r = rc0 + j;
            Yt[l + j * b] = ldexpf(orc_e4m3_to_f32(Rp->q[r + l * Rp->rows]), -Rp->exps[r / tile + (l / tile) * rrt]);
        }
    gemm_seq_core(mr, nc, b, -1.0f, Xf, Yt, 1.0f, C, ldc);
    free(Xf);
    free(Yt);
}

/* Blocked right-looking no-pivot LU in the step order of the reference's
 * default driver panel_lu_async_wGPU_gdirect_alt2 on a 1x1 grid
 * (panel_trf_gdirect.hpp:699-737 step 0, :777-915 steps k >= 1):
 *   step 0: getrf(A00); A[1:,0] <- A[1:,0] U00^-1 and A[0,1:] <- L00^-1 A[0,1:]
 *           by FP32 TRSM (panel_trf.hpp:188-246); pack both panels.
 *   step k: A_kk -= L_{k-1}[k] R_{k-1}[k]^T; getrf(A_kk); piv = copy(A_kk);
 *           alt 2: piv <- trtri(U), trtri(L) in place;
 *           A[k+1:,k] -= L_{k-1}[k+1:] R_{k-1}[k]^T; column panel by alt;
 *           pack L_k; A[k,k+1:] -= L_{k-1}[k] R_{k-1}[k+1:]^T; row panel by
 *           alt; pack R_k; A[k+1:,k+1:] -= L_{k-1}[k+1:] R_{k-1}[k+1:]^T.
 * Column panel by alt (:844-859, :46-163):
 *   0: FP32 TRSM with U_kk;
 *   1: W = U_kk^-1 by TRSM on the identity; A = f16(A) * f16(W) (FP32 acc);
 *   2: W = triu(trtri(U_kk)) (lower part zero); A = f16(A) * f16(W).
 * Row panel by alt (:876-891, :110-142, :165-183):
 *   0: FP32 TRSM with unit L_kk;
 *   1: W = L_kk^-1 (X L = I on the identity); A = f16(W) * f16(A);
 *   2: W = strict tril(trtri(L_kk)) + I; A = f16(W) * f16(A).
 * The Schur panels are FP16 (reference) or, where sched[k] = 1, FP8 e4m3
 * with power-of-two scales per fp8_tile x fp8_tile tile. */
int orc_lu_factor(int64_t n, int64_t b, int alt, float *A, int64_t lda,
                  const int *sched, int64_t fp8_tile, double *stats) {
    if (n <= 0 || b <= 0 || n % b) return -1;
    const int64_t nb = n / b;
    if (fp8_tile <= 0) fp8_tile = 256;
    int info = 0;
    orc_panel_t Lprev, Rprev, Lnext, Rnext;
    const int64_t rmax = n > b ? n - b : 1;
    panel_alloc(&Lprev, rmax, b, fp8_tile);
    panel_alloc(&Rprev, rmax, b, fp8_tile);
    panel_alloc(&Lnext, rmax, b, fp8_tile);
    panel_alloc(&Rnext, rmax, b, fp8_tile);
    float *piv = (float *)malloc(sizeof(float) * (size_t)(b * b));
    float *W32 = (float *)malloc(sizeof(float) * (size_t)(b * b));
    uint16_t *Wh = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)(b * b));
    uint16_t *Ph = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)(rmax * b + 1));

#define AT(i, j) (A + (i) + (j) * lda)
    for (int64_t k = 0; k < nb; ++k) {
        const int64_t kb = k * b;
        const int64_t rest = n - kb - b; /* rows/cols after block k */
        float *Akk = AT(kb, kb);
        /* the step's Schur class: 1 FP8, -1 FP32, 0 FP16 */
        int fp8 = (sched && sched[k]) ? (sched[k] < 0 ? -1 : 1) : 0;
        if (k > 0) {
            /* diagonal block update with the previous panels (:791-795) */
            panel_schur(&Lprev, 0, &Rprev, 0, b, b, Akk, lda);
        }
        const int inf = orc_getrf_nopiv(b, Akk, lda);
        if (inf && !info) info = (int)(kb + inf);
        double mindiag = INFINITY;
        for (int64_t i = 0; i < b; ++i) {
            const double v = fabs((double)Akk[i + i * lda]);
            if (v < mindiag) mindiag = v;
        }
        for (int64_t j = 0; j < b; ++j) memcpy(piv + j * b, Akk + j * lda, sizeof(float) * (size_t)b);
        if (k > 0 && alt == 2) {
            orc_trtri_upper_nonunit(b, piv, b);
            orc_trtri_lower_unit(b, piv, b);
        }
        double mxl = 0.0, mxr = 0.0;
        if (rest > 0) {
            float *Acol = AT(kb + b, kb);
            float *Arow = AT(kb, kb + b);
            /* column panel */
            if (k > 0) panel_schur(&Lprev, b, &Rprev, 0, rest, b, Acol, lda);
            if (k == 0 || alt == 0 || fp8 < 0) {
                /* the factored block itself (piv is its copy for -alt 0, its
                 * inverses for -alt 2) */
                orc_trsm_runn(rest, b, Akk, lda, Acol, lda);
            } else {
                /* W = U^-1, lower part zero */
                if (alt == 1) {
                    for (int64_t j = 0; j < b; ++j)
                        for (int64_t i = 0; i < b; ++i) W32[i + j * b] = (i == j) ? 1.0f : 0.0f;
                    orc_trsm_runn(b, b, piv, b, W32, b);
                    for (int64_t i = 0; i < b * b; ++i) Wh[i] = orc_f32_to_f16(W32[i]);
                } else {
                    for (int64_t j = 0; j < b; ++j)
                        for (int64_t i = 0; i < b; ++i)
                            Wh[i + j * b] = (i <= j) ? orc_f32_to_f16(piv[i + j * b]) : 0;
                }
                orc_cast_f16(rest, b, Acol, lda, Ph, rest);
                /* Acol = Ph * W : X = Ph (rest x b), Y(j,l) = W(l,j) */
                {
                    uint16_t *Wt = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)(b * b));
                    for (int64_t j = 0; j < b; ++j)
                        for (int64_t l = 0; l < b; ++l) Wt[j + l * b] = Wh[l + j * b];
                    orc_gemm_f16(rest, b, b, 1.0f, Ph, rest, Wt, b, 0.0f, Acol, lda);
                    free(Wt);
                }
            }
            panel_fill(&Lnext, rest, Acol, 1, lda, fp8, &mxl);

            /* row panel */
            if (k > 0) panel_schur(&Lprev, 0, &Rprev, b, b, rest, Arow, lda);
            if (k == 0 || alt == 0 || fp8 < 0) {
                orc_trsm_llnu(b, rest, Akk, lda, Arow, lda);
            } else {
                if (alt == 1) {
                    for (int64_t j = 0; j < b; ++j)
                        for (int64_t i = 0; i < b; ++i) W32[i + j * b] = (i == j) ? 1.0f : 0.0f;
                    orc_trsm_rlnu(b, b, piv, b, W32, b);
                    for (int64_t i = 0; i < b * b; ++i) Wh[i] = orc_f32_to_f16(W32[i]);
                } else {
                    for (int64_t j = 0; j < b; ++j)
                        for (int64_t i = 0; i < b; ++i)
                            Wh[i + j * b] = (i > j) ? orc_f32_to_f16(piv[i + j * b]) : (i == j ? 0x3c00 : 0);
                }
                /* Ph(j, l) = f16(Arow(l, j)) : transposed row panel */
#pragma omp parallel for schedule(static)
                for (int64_t l = 0; l < b; ++l)
                    for (int64_t j = 0; j < rest; ++j) Ph[j + l * rest] = orc_f32_to_f16(Arow[l + j * lda]);
                /* Arow = W * Ph^T : X = W (b x b), Y = Ph (rest x b) */
                orc_gemm_f16(b, rest, b, 1.0f, Wh, b, Ph, rest, 0.0f, Arow, lda);
            }
            panel_fill(&Rnext, rest, Arow, lda, 1, fp8, &mxr);

            /* trailing update with the previous panels (:900-915) */
            if (k > 0) panel_schur(&Lprev, b, &Rprev, b, rest, rest, AT(kb + b, kb + b), lda);
        }
        if (stats) {
            stats[4 * k + 0] = mxl;
            stats[4 * k + 1] = mxr;
            stats[4 * k + 2] = 0.0;
            stats[4 * k + 3] = mindiag;
        }
        orc_panel_t t = Lprev;
        Lprev = Lnext;
        Lnext = t;
        t = Rprev;
        Rprev = Rnext;
        Rnext = t;
        /* step 0 leaves its trailing update to step 1 (look-ahead) */
        if (k == 0 && rest > 0) {
            /* nothing: the reference applies step 0's panels inside step 1 */
        }
    }
#undef AT
    panel_free(&Lprev);
    panel_free(&Rprev);
    panel_free(&Lnext);
    panel_free(&Rnext);
    free(piv);
    free(W32);
    free(Wh);
    free(Ph);
    return info;
}

/* ------------------------------------------------------------------------ */
/* refinement                                                                */
/* ------------------------------------------------------------------------ */

/* y <- y + alpha * A * x with A regenerated from the LCG and A_ii = diag[i]
 * (otf_gemv_d / otf_gemv_diag_d, panel_gemv.hpp:13-117).  Per row the sum
 * runs over columns in order in FP64. */
void orc_otf_gemv(int64_t n, uint64_t seed, const double *diag, double alpha,
                  const double *x, double *y) {
    const uint64_t x00 = orc_lcg_init(seed);
    const orc_map_t jump_n = map_pow((uint64_t)n);
    enum { RB = 64 };
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t r0 = 0; r0 < n; r0 += RB) {
        const int64_t r1 = (r0 + RB < n) ? r0 + RB : n;
        double acc[RB];
        for (int i = 0; i < RB; ++i) acc[i] = 0.0;
        uint64_t st = map_apply(map_pow((uint64_t)r0), x00);
        for (int64_t j = 0; j < n; ++j) {
            const double xj = x[j];
            uint64_t s = st;
            for (int64_t i = r0; i < r1; ++i) {
                const double a = (i == j) ? diag[i] : orc_lcg_value(s);
                acc[i - r0] += a * xj;
                s = ORC_LCG_A * s + ORC_LCG_C;
            }
            st = map_apply(jump_n, st);
        }
        for (int64_t i = r0; i < r1; ++i) y[i] = y[i] + alpha * acc[i - r0];
    }
}

/* colv_infnorm (panel_norm.hpp:11-90): max |x_i| */
double orc_norm_inf(int64_t n, const double *x) {
    double m = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double t = fabs(x[i]);
        m = (m >= t) ? m : t;
    }
    return m;
}

/* hpl-harness (iterative_refinement.hpp:154-157) */
double orc_harness(double normr, double norma, double normx, double normb,
                   int64_t n) {
    return normr / (norma * normx + normb) * 1.0 / ((double)n * DBL_EPSILON / 2);
}

/* x <- L^-1 x with the unit-lower FP32 factor, FP64 arithmetic, blocked by
 * b: the diagonal block is solved in FP64 on the upcast factor
 * (cast_copy_general + rocblas_dtrsv) after subtracting the accumulated
 * contributions; the sub-diagonal blocks add L_ij x_j into an FP64
 * accumulator (trsvL_h, panel_trsv.hpp:308-449, with the 1x1 ring swap of
 * the accumulation buffers not reproduced). */
void orc_trsv_lower_unit(int64_t n, int64_t b, const float *A, int64_t lda,
                         double *x) {
    double *acc = (double *)calloc((size_t)n, sizeof(double));
    const int64_t nb = n / b;
    for (int64_t jb = 0; jb < nb; ++jb) {
        const int64_t j0 = jb * b;
        double *xj = x + j0;
        for (int64_t i = 0; i < b; ++i) xj[i] -= acc[j0 + i];
        for (int64_t l = 0; l < b; ++l) {
            const double t = xj[l];
            const float *c = A + j0 + (j0 + l) * lda;
            for (int64_t i = l + 1; i < b; ++i) xj[i] -= t * (double)c[i];
        }
        const int64_t r0 = j0 + b;
#pragma omp parallel for schedule(static)
        for (int64_t i0 = r0; i0 < n; i0 += 256) {
            const int64_t i1 = (i0 + 256 < n) ? i0 + 256 : n;
            double w[256];
            for (int64_t i = i0; i < i1; ++i) w[i - i0] = 0.0;
            for (int64_t l = 0; l < b; ++l) {
                const double t = xj[l];
                const float *c = A + (j0 + l) * lda;
                for (int64_t i = i0; i < i1; ++i) w[i - i0] += (double)c[i] * t;
            }
            for (int64_t i = i0; i < i1; ++i) acc[i] += w[i - i0];
        }
    }
    free(acc);
}

/* x <- U^-1 x with the upper non-unit FP32 factor (trsvU_h,
 * panel_trsv.hpp:452-597), block columns from the last to the first. */
void orc_trsv_upper(int64_t n, int64_t b, const float *A, int64_t lda,
                    double *x) {
    double *acc = (double *)calloc((size_t)n, sizeof(double));
    const int64_t nb = n / b;
    for (int64_t jb = nb - 1; jb >= 0; --jb) {
        const int64_t j0 = jb * b;
        double *xj = x + j0;
        for (int64_t i = 0; i < b; ++i) xj[i] -= acc[j0 + i];
        for (int64_t l = b - 1; l >= 0; --l) {
            const float *c = A + j0 + (j0 + l) * lda;
            xj[l] /= (double)c[l];
            const double t = xj[l];
            for (int64_t i = 0; i < l; ++i) xj[i] -= t * (double)c[i];
        }
#pragma omp parallel for schedule(static)
        for (int64_t i0 = 0; i0 < j0; i0 += 256) {
            const int64_t i1 = (i0 + 256 < j0) ? i0 + 256 : j0;
            double w[256];
            for (int64_t i = i0; i < i1; ++i) w[i - i0] = 0.0;
            for (int64_t l = 0; l < b; ++l) {
                const double t = xj[l];
                const float *c = A + (j0 + l) * lda;
                for (int64_t i = i0; i < i1; ++i) w[i - i0] += (double)c[i] * t;
            }
            for (int64_t i = i0; i < i1; ++i) acc[i] += w[i - i0];
        }
    }
    free(acc);
}

/* Classical IR, GPU branch semantics (iterative_refinement.hpp:141-171):
 * x0 = b ./ d; per step r = b - A x (FP64, A regenerated), harness check,
 * then x += U^-1 L^-1 r.  norma = 2 max d (panel_norm.hpp:129-135). */
int orc_ir_classic(int64_t n, int64_t b, uint64_t seed, const float *LU,
                   int64_t lda, const double *diag, const double *rhs,
                   int maxit, double *x, double *resid, double *harness,
                   double *trace) {
    double *r = (double *)malloc(sizeof(double) * (size_t)n);
    const double norma = 2.0 * orc_norm_inf(n, diag);
    const double normb = orc_norm_inf(n, rhs);
    for (int64_t i = 0; i < n; ++i) x[i] = rhs[i] / diag[i];
    int result = -1;
    *resid = -1.0;
    *harness = -1.0;
    for (int it = 0; it < maxit; ++it) {
        memcpy(r, rhs, sizeof(double) * (size_t)n);
        const double normx = orc_norm_inf(n, x);
        orc_otf_gemv(n, seed, diag, -1.0, x, r);
        const double normr = orc_norm_inf(n, r);
        const double h = orc_harness(normr, norma, normx, normb, n);
        if (trace) {
            trace[2 * it] = normr;
            trace[2 * it + 1] = h;
        }
        if (h < 16.0) {
            *resid = normr;
            *harness = h;
            result = it;
            break;
        }
        orc_trsv_lower_unit(n, b, LU, lda, r);
        orc_trsv_upper(n, b, LU, lda, r);
        for (int64_t i = 0; i < n; ++i) x[i] += r[i];
    }
    free(r);
    return result;
}

static double dot(int64_t n, const double *a, const double *b) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static void precond(int64_t n, int64_t b, const float *LU, int64_t lda, double *v) {
    orc_trsv_lower_unit(n, b, LU, lda, v);
    orc_trsv_upper(n, b, LU, lda, v);
}

/* GMRES-IR: the outer loop is the classical one (same x0, residual and
 * harness); each correction solves A d = r by GMRES(restart) on the
 * left-preconditioned system (LU)^-1 A d = (LU)^-1 r, modified Gram-Schmidt
 * Arnoldi and Givens rotations, all in FP64, stopping when the
 * preconditioned residual falls by inner_tol. */
int orc_gmres_ir(int64_t n, int64_t b, uint64_t seed, const float *LU,
                 int64_t lda, const double *diag, const double *rhs,
                 int maxit, int restart, double inner_tol, double *x,
                 double *resid, double *harness, double *trace,
                 int *inner_total) {
    const int m = restart;
    double *r = (double *)malloc(sizeof(double) * (size_t)n);
    double *V = (double *)malloc(sizeof(double) * (size_t)n * (size_t)(m + 1));
    double *w = (double *)malloc(sizeof(double) * (size_t)n);
    double *H = (double *)calloc((size_t)(m + 1) * (size_t)m, sizeof(double));
    double *cs = (double *)malloc(sizeof(double) * (size_t)m);
    double *sn = (double *)malloc(sizeof(double) * (size_t)m);
    double *g = (double *)malloc(sizeof(double) * (size_t)(m + 1));
    double *yv = (double *)malloc(sizeof(double) * (size_t)m);
    const double norma = 2.0 * orc_norm_inf(n, diag);
    const double normb = orc_norm_inf(n, rhs);
    for (int64_t i = 0; i < n; ++i) x[i] = rhs[i] / diag[i];
    int result = -1, inner = 0;
    *resid = -1.0;
    *harness = -1.0;
    for (int it = 0; it < maxit; ++it) {
        memcpy(r, rhs, sizeof(double) * (size_t)n);
        const double normx = orc_norm_inf(n, x);
        orc_otf_gemv(n, seed, diag, -1.0, x, r);
        const double normr = orc_norm_inf(n, r);
        const double h = orc_harness(normr, norma, normx, normb, n);
        if (trace) {
            trace[2 * it] = normr;
            trace[2 * it + 1] = h;
        }
        if (h < 16.0) {
            *resid = normr;
            *harness = h;
            result = it;
            break;
        }
        /* GMRES on (LU)^-1 A d = (LU)^-1 r */
        double *v0 = V;
        memcpy(v0, r, sizeof(double) * (size_t)n);
        precond(n, b, LU, lda, v0);
        const double beta = sqrt(dot(n, v0, v0));
        if (!(beta > 0.0)) break;
        for (int64_t i = 0; i < n; ++i) v0[i] /= beta;
        for (int i = 0; i <= m; ++i) g[i] = 0.0;
        g[0] = beta;
        int jdone = 0;
        for (int j = 0; j < m; ++j) {
            double *vj = V + (size_t)j * (size_t)n;
            for (int64_t i = 0; i < n; ++i) w[i] = 0.0;
            orc_otf_gemv(n, seed, diag, 1.0, vj, w);
            precond(n, b, LU, lda, w);
            for (int i = 0; i <= j; ++i) {
                const double *vi = V + (size_t)i * (size_t)n;
                const double hij = dot(n, w, vi);
                H[i + j * (m + 1)] = hij;
                for (int64_t q = 0; q < n; ++q) w[q] -= hij * vi[q];
            }
            const double hn = sqrt(dot(n, w, w));
            H[(j + 1) + j * (m + 1)] = hn;
            double *vn = V + (size_t)(j + 1) * (size_t)n;
            for (int64_t q = 0; q < n; ++q) vn[q] = (hn > 0.0) ? w[q] / hn : 0.0;
            for (int i = 0; i < j; ++i) {
                const double a = H[i + j * (m + 1)], bb = H[(i + 1) + j * (m + 1)];
                H[i + j * (m + 1)] = cs[i] * a + sn[i] * bb;
                H[(i + 1) + j * (m + 1)] = -sn[i] * a + cs[i] * bb;
            }
            const double a = H[j + j * (m + 1)], bb = H[(j + 1) + j * (m + 1)];
            const double den = hypot(a, bb);
            cs[j] = (den > 0.0) ? a / den : 1.0;
            sn[j] = (den > 0.0) ? bb / den : 0.0;
            H[j + j * (m + 1)] = den;
            H[(j + 1) + j * (m + 1)] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            jdone = j + 1;
            ++inner;
            if (fabs(g[j + 1]) <= inner_tol * beta || hn == 0.0) break;
        }
        for (int i = jdone - 1; i >= 0; --i) {
            double s = g[i];
            for (int q = i + 1; q < jdone; ++q) s -= H[i + q * (m + 1)] * yv[q];
            yv[i] = s / H[i + i * (m + 1)];
        }
        for (int i = 0; i < jdone; ++i) {
            const double *vi = V + (size_t)i * (size_t)n;
            for (int64_t q = 0; q < n; ++q) x[q] += yv[i] * vi[q];
        }
    }
    if (inner_total) *inner_total = inner;
    free(r);
    free(V);
    free(w);
    free(H);
    free(cs);
    free(sn);
    free(g);
    free(yv);
    return result;
}

/* ------------------------------------------------------------------------ */
/* randsvd test matrices (new; no reference counterpart)                     */
/* ------------------------------------------------------------------------ */
static void randsvd_reflector(int64_t n, uint64_t seed, int t, double *w) {
    double nrm2 = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        /* linear index n^2 + n + t*n + i = i + n*(n + 1 + t) */
        w[i] = orc_lcg_value(orc_lcg_state(seed, (uint64_t)n, (uint64_t)i, (uint64_t)(n + 1 + t)));
        nrm2 += w[i] * w[i];
    }
    const double inv = 1.0 / sqrt(nrm2);
    for (int64_t i = 0; i < n; ++i) w[i] *= inv;
}

void orc_randsvd_dense(int64_t n, uint64_t seed, double kappa, int nrefl,
                       double *A) {
    double *w = (double *)malloc(sizeof(double) * (size_t)n);
    double *t = (double *)malloc(sizeof(double) * (size_t)n);
    memset(A, 0, sizeof(double) * (size_t)n * (size_t)n);
    for (int64_t i = 0; i < n; ++i)
        A[i + i * n] = (n == 1) ? 1.0 : pow(kappa, -(double)i / (double)(n - 1));
    /* right factors: A <- A (I - 2 w w^T), t = 2k-1 .. k */
    for (int r = 2 * nrefl - 1; r >= nrefl; --r) {
        randsvd_reflector(n, seed, r, w);
        for (int64_t i = 0; i < n; ++i) {
            double s = 0.0;
            for (int64_t j = 0; j < n; ++j) s += A[i + j * n] * w[j];
            t[i] = s;
        }
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < n; ++i) A[i + j * n] -= 2.0 * t[i] * w[j];
    }
    /* left factors: A <- (I - 2 w w^T) A, t = k-1 .. 0 */
    for (int r = nrefl - 1; r >= 0; --r) {
        randsvd_reflector(n, seed, r, w);
        for (int64_t j = 0; j < n; ++j) {
            double s = 0.0;
            for (int64_t i = 0; i < n; ++i) s += w[i] * A[i + j * n];
            t[j] = s;
        }
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < n; ++i) A[i + j * n] -= 2.0 * w[i] * t[j];
    }
    free(w);
    free(t);
}

/* ------------------------------------------------------------------------ */
/* MLP 16-64-64-3 (new; the learned FP8/FP16 schedule)                        */
/* ------------------------------------------------------------------------ */
#define W1_OFF 0
#define B1_OFF (W1_OFF + ORC_MLP_IN * ORC_MLP_H)
#define W2_OFF (B1_OFF + ORC_MLP_H)
#define B2_OFF (W2_OFF + ORC_MLP_H * ORC_MLP_H)
#define W3_OFF (B2_OFF + ORC_MLP_H)
#define B3_OFF (W3_OFF + ORC_MLP_H * ORC_MLP_OUT)

static void mlp_fwd1(const float *p, const float *x, float *h1, float *h2, float *z) {
    for (int j = 0; j < ORC_MLP_H; ++j) {
        float a = p[B1_OFF + j];
        for (int i = 0; i < ORC_MLP_IN; ++i) a += x[i] * p[W1_OFF + i * ORC_MLP_H + j];
        h1[j] = a > 0.0f ? a : 0.0f;
    }
    for (int j = 0; j < ORC_MLP_H; ++j) {
        float a = p[B2_OFF + j];
        for (int i = 0; i < ORC_MLP_H; ++i) a += h1[i] * p[W2_OFF + i * ORC_MLP_H + j];
        h2[j] = a > 0.0f ? a : 0.0f;
    }
    for (int j = 0; j < ORC_MLP_OUT; ++j) {
        float a = p[B3_OFF + j];
        for (int i = 0; i < ORC_MLP_H; ++i) a += h2[i] * p[W3_OFF + i * ORC_MLP_OUT + j];
        z[j] = a;
    }
}

void orc_mlp_forward(int64_t batch, const float *params, const float *X,
                     float *H1, float *H2, float *Z) {
    for (int64_t s = 0; s < batch; ++s)
        mlp_fwd1(params, X + s * ORC_MLP_IN, H1 + s * ORC_MLP_H, H2 + s * ORC_MLP_H,
                 Z + s * ORC_MLP_OUT);
}

float orc_mlp_backward(int64_t batch, const float *params, const float *X,
                       const int32_t *y, const float *t, const float *m,
                       float *grad) {
    float h1[ORC_MLP_H], h2[ORC_MLP_H], z[ORC_MLP_OUT];
    float dz[ORC_MLP_OUT], d2[ORC_MLP_H], d1[ORC_MLP_H];
    double loss = 0.0;
    memset(grad, 0, sizeof(float) * ORC_MLP_NPARAMS);
    const float inv = 1.0f / (float)batch;
    for (int64_t s = 0; s < batch; ++s) {
        const float *x = X + s * ORC_MLP_IN;
        mlp_fwd1(params, x, h1, h2, z);
        /* softmax cross-entropy over (z0, z1) */
        const float mx = z[0] > z[1] ? z[0] : z[1];
        const float e0 = expf(z[0] - mx), e1 = expf(z[1] - mx);
        const float lse = mx + logf(e0 + e1);
        const float p0 = e0 / (e0 + e1), p1 = e1 / (e0 + e1);
        const float diff = z[2] - t[s];
        loss += (double)(lse - (y[s] ? z[1] : z[0])) + 0.5 * (double)(m[s] * diff * diff);
        dz[0] = (p0 - (y[s] ? 0.0f : 1.0f)) * inv;
        dz[1] = (p1 - (y[s] ? 1.0f : 0.0f)) * inv;
        dz[2] = m[s] * diff * inv;
        for (int j = 0; j < ORC_MLP_OUT; ++j) grad[B3_OFF + j] += dz[j];
        for (int i = 0; i < ORC_MLP_H; ++i) {
            float a = 0.0f;
            for (int j = 0; j < ORC_MLP_OUT; ++j) {
                grad[W3_OFF + i * ORC_MLP_OUT + j] += h2[i] * dz[j];
                a += params[W3_OFF + i * ORC_MLP_OUT + j] * dz[j];
            }
            d2[i] = h2[i] > 0.0f ? a : 0.0f;
        }
        for (int j = 0; j < ORC_MLP_H; ++j) grad[B2_OFF + j] += d2[j];
        for (int i = 0; i < ORC_MLP_H; ++i) {
            float a = 0.0f;
            for (int j = 0; j < ORC_MLP_H; ++j) {
                grad[W2_OFF + i * ORC_MLP_H + j] += h1[i] * d2[j];
                a += params[W2_OFF + i * ORC_MLP_H + j] * d2[j];
            }
            d1[i] = h1[i] > 0.0f ? a : 0.0f;
        }
        for (int j = 0; j < ORC_MLP_H; ++j) grad[B1_OFF + j] += d1[j];
        for (int i = 0; i < ORC_MLP_IN; ++i)
            for (int j = 0; j < ORC_MLP_H; ++j) grad[W1_OFF + i * ORC_MLP_H + j] += x[i] * d1[j];
    }
    return (float)(loss / (double)batch);
}
