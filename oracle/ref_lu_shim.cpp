// ref_lu_shim.cpp -- compiles the reference's sgetrf_nopiv.cpp as it lies
// (included from the reference tree by the Makefile's include path) so that
// its internal-linkage sgetrf_nopiv_naive can be called by the oracle tests.
//
// TEST INFRASTRUCTURE ONLY: built into oracle/_ref/ by oracle/Makefile when
// the reference sources are present; loaded only by tests/.  Nothing here
// restates an algorithm: the wrapper forwards to the reference routine.
#include "sgetrf_nopiv.cpp"  // sgetrf_nopiv.cpp:27-61, :938-940

extern "C" {

// sgetrf_nopiv_naive (sgetrf_nopiv.cpp:27-61) on an m x n block
void ref_sgetrf_nopiv_naive(int m, int n, float *a, size_t lda) {
    sgetrf_nopiv_naive(m, n, a, lda);
}

}  // extern "C"
