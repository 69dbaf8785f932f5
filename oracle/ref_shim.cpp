// ref_shim.cpp -- exposes host routines of the reference (at-aaims/OpenMxP),
// compiled from its own sources where they lie, to the oracle tests.
//
// TEST INFRASTRUCTURE ONLY: built into oracle/_ref/ by oracle/Makefile when
// the reference sources are present; loaded only by tests/ to pin the CPU
// oracle.  Nothing here restates an algorithm: each function forwards to the
// reference routine named in its comment.
#include <cstddef>
#include <cstdint>

#include "hpl_rand.hpp"  // hpl_rand.hpp:12-255
#include "fp16sim.hpp"   // fp16sim.hpp (clang: FP16_AUTO_PROMOTION path)

// hpl_rand.hpp:10 declares `extern int grank`; main.cpp:31 defines it.
int grank = 0;

// sgetrf_nopiv.cpp:938-940 (OpenMP column-cyclic no-pivot LU, n <= 1280),
// compiled from the reference file into the same .so
void sgetrf_nopiv_tuned(int n, float *a, size_t lda);

// higham_mat_impl.cpp:36-64, compiled from the reference file into the
// same .so (the extern "C" routines of that file, higham_mat_comp_beta and
// hplai_matrix_impl, are exported as they are)
double fhpl(int n, double alpha, double beta);

extern "C" {

// RandStat of linear index i + n*j, as double (hpl_rand.hpp:54-65, :249-254)
double ref_lcg_value(uint64_t seed, int n, uint64_t i, uint64_t j) {
    RandStat s = RandStat::initialize(seed);
    RandCoeff c = RandCoeff::default_vals().pow(i + static_cast<uint64_t>(n) * j);
    return static_cast<double>(c * s);
}

// RandCoeff::pow (hpl_rand.hpp:31-48)
void ref_lcg_jump(uint64_t steps, uint64_t *a, uint64_t *c) {
    RandCoeff r = RandCoeff::default_vals().pow(steps);
    *a = r.a;
    *c = r.c;
}

// RandCoeff::pow_fugaku (hpl_rand.hpp:21-30), a second powering routine
void ref_lcg_jump_fugaku(uint64_t steps, uint64_t *a, uint64_t *c) {
    RandCoeff r = RandCoeff::default_vals().pow_fugaku(steps);
    *a = r.a;
    *c = r.c;
}

// panel_fill_one_with_rand without the diagonal (hpl_rand.hpp:80-110)
void ref_fill_block(int n, int i0, int j0, int nrow, int ncol, float *a,
                    size_t lda, uint64_t seed) {
    panel_fill_one_with_rand<float>(n, i0, j0, nrow, ncol, a, lda, seed, false);
}

// panel_fill_one_with_rand with calc_diag (hpl_rand.hpp:80-129): diagonal
// block with A_ii = float(sequential row sum)
void ref_fill_block_diag(int n, int i0, int nrow, float *a, size_t lda,
                         uint64_t seed) {
    panel_fill_one_with_rand<float>(n, i0, i0, nrow, nrow, a, lda, seed, true);
}

// calc_diag (hpl_rand.hpp:200-214)
double ref_calc_diag(int i, int n, uint64_t seed) { return calc_diag(i, n, seed); }

// right-hand side as panel_colvgen builds it (matgen.hpp:433)
void ref_rhs(int n, int i0, int m, double *b, uint64_t seed) {
    panel_fill_one_with_rand<double>(n, i0, n, m, 1, b, 1, seed, false);
}

// fp16 conversion of fp16sim.hpp (FP16_AUTO_PROMOTION: __fp16)
uint16_t ref_f32_to_f16(float f) {
    fp16 h = f;
    uint16_t bits;
    __builtin_memcpy(&bits, &h, 2);
    return bits;
}

float ref_f16_to_f32(uint16_t bits) {
    fp16 h;
    __builtin_memcpy(&h, &bits, 2);
    return static_cast<float>(h);
}

// shgemm (fp16sim.hpp:84-99): c = c*beta + alpha * sum_l a(j,l) b(l,i)
void ref_shgemm(int m, int n, int k, float alpha, const uint16_t *a, int lda,
                const uint16_t *b, int ldb, float beta, float *c, int ldc) {
    shgemm(m, n, k, alpha, reinterpret_cast<const fp16 *>(a), lda,
           reinterpret_cast<const fp16 *>(b), ldb, beta, c, ldc);
}

// sgetrf_nopiv_tuned (sgetrf_nopiv.cpp:838-940); blocks of 6 columns whose
// panels use the naive kernel of sgetrf_nopiv.cpp:27-61
void ref_sgetrf_nopiv_tuned(int n, float *a, size_t lda) {
    sgetrf_nopiv_tuned(n, a, lda);
}

// fhpl (higham_mat_impl.cpp:36-64): the cond_inf estimate of A(n, a, b)
double ref_higham_fhpl(int n, double alpha, double beta) { return fhpl(n, alpha, beta); }

}  // extern "C"
