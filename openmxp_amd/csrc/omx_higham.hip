// omx_higham.hip -- the Fasi-Higham HPL-AI test matrix (the reference's
// --highammat family, highammgen.hpp:1-31, higham_mat_impl.cpp:37-160):
//   A = L U,  L unit lower with -alpha below the diagonal,
//             U unit upper with -beta above it,
// i.e. a_ij = -beta + (alpha beta) i above the diagonal, 1 + (alpha beta) i
// on it, -alpha + (alpha beta) j below it (0-based).  The generator writes
// the FP32 rounding of these FP64 values; the refinement's residual
// regenerates the FP64 values row by row.
#include <float.h>

#include <algorithm>
#include <cmath>

#include "omx_common.hpp"

// entries are evaluated as written: the product, then the sum (no fused
// multiply-add), so that host and device values agree bit for bit
#pragma clang fp contract(off)

namespace {

__host__ __device__ __forceinline__ double higham_entry(int64_t i, int64_t j, double alpha, double beta, double ab) {
    if (i < j) return -beta + ab * (double)i;
    if (i == j) return 1.0 + ab * (double)i;
    return -alpha + ab * (double)j;
}

// One workgroup per column (grid-stride), a thread per row (stride 256):
// every wave stores a contiguous run of the column.
__global__ void __launch_bounds__(256)
k_higham_fill(int64_t n, double alpha, double beta, float *__restrict__ A, int64_t lda, double *__restrict__ diag) {
    const double ab = alpha * beta;
    for (int64_t j = blockIdx.x; j < n; j += gridDim.x) {
        float *col = A + j * lda;
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
            const double v = higham_entry(i, j, alpha, beta, ab);
            col[i] = (float)v;
            if (i == j) diag[j] = v;
        }
    }
}

// y_i += s sum_j a_ij x_j, a thread per row, columns in order with one FMA
// each (x_j is the same for the whole wave: a scalar load).
__global__ void __launch_bounds__(256)
k_higham_gemv(int64_t n, double alpha, double beta, double s, const double *__restrict__ x,
              double *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double ab = alpha * beta;
    const double up = -beta + ab * (double)i;  // every entry right of the diagonal
    double acc = 0.0;
    for (int64_t j = 0; j < i; ++j) acc = __builtin_fma(-alpha + ab * (double)j, x[j], acc);
    acc = __builtin_fma(1.0 + ab * (double)i, x[i], acc);
    for (int64_t j = i + 1; j < n; ++j) acc = __builtin_fma(up, x[j], acc);
    y[i] = __builtin_fma(s, acc, y[i]);
}

// The paper's estimate of cond_inf(A(n, a, b)): an estimate of ||A||_inf
// (the largest of three row sums of the closed form) times one of
// ||A^-1||_inf (the larger of the first and last row sums of the inverse).
// The reference evaluates the integer products in 32 bits, which overflows
// for n above 46340; 64 bits here.
double cond_estimate(int64_t n, double a, double b) {
    if (std::isnan(a)) a = b / 2;
    const double l1 = 1 + (double)(n - 1) * b;
    const int64_t id = std::min<int64_t>((int64_t)std::floor(1.0 / a), n);
    const int64_t k = std::min<int64_t>((int64_t)std::floor((1 + b) / b), n - 1);
    const double lid = 1 + (double)(2 * k - id + 1) * a + (double)(n - id) * b +
                       (double)(-k * k + k + 3 * id * (id - 1) / 2 - n * id + n) * a * b;
    const double ln = 1 + (double)(2 * k - n + 1) * a + (double)(-k * k + k + n * (n - 1) / 2) * a * b;
    const double norm_a = std::max(std::max(l1, lid), ln);
    const double r = (1 + a) * (1 + b);
    const double d1 = (1 + a) * (1.0 / (1 + a) + (r == 0.0 ? 0.0 : b * (1 - std::pow(r, (double)(n - 1))) / (1 - r)));
    const double dn = std::pow(1 + a, (double)n) * (1.0 / (1 + a));
    const double v = norm_a * std::max(d1, dn);
    return std::isinf(v) ? DBL_MAX : v;
}

// Bisection for f(x) = 0 on [lo, hi] with f(lo) <= 0 <= f(hi), stopped when
// the interval is below (2 |hi| + 1/2) eps or its midpoint stops moving;
// returns the upper end (NaN without a sign change).
template <typename F>
double bisect(F f, double lo, double hi) {
    if (f(lo) > 0.0 || f(hi) < 0.0) return std::nan("");
    for (;;) {
        if (hi - lo < (2.0 * std::fabs(hi) + 0.5) * DBL_EPSILON) break;
        const double mid = (lo + hi) / 2;
        if (mid == lo || mid == hi) break;
        const double fm = f(mid);
        if (fm == 0.0) return mid;
        if (fm < 0.0) lo = mid;
        else hi = mid;
    }
    return hi;
}

}  // namespace

extern "C" {

double omx_higham_cond_est(int64_t n, double alpha, double beta) {
    if (n < 2) return std::nan("");
    return cond_estimate(n, alpha, beta);
}

double omx_higham_beta(int64_t n, double kappa, double rho) {
    if (n < 2 || !(kappa > 1.0) || !(rho > 0.0)) return std::nan("");
    auto f = [&](double x) { return cond_estimate(n, rho * x, x) - kappa; };
    // the left end: the smallest beta; the right end from 1/rho, halved
    // until the estimate is finite and above kappa (at most 100 times)
    const double lo = DBL_EPSILON;
    double hi = 1.0 / rho;
    for (int k = 1; k < 100; ++k) {
        const double v = f(hi);
        if (std::isfinite(v) && v > 0.0) break;
        hi *= 0.5;
    }
    double beta = bisect(f, lo, hi);
    // alpha must not exceed 1: shrink the bracket by four and search again
    while (rho * beta > 1.0) {
        hi *= 0.25;
        beta = bisect(f, lo, hi);
    }
    return beta;
}

double omx_higham_norm_inf(int64_t n, double alpha, double beta) {
    if (n <= 0) return std::nan("");
    const double ab = alpha * beta;
    double left = 0.0, best = 0.0;  // left = sum_{j < i} |a_ij|
    for (int64_t i = 0; i < n; ++i) {
        const double row = left + std::fabs(higham_entry(i, i, alpha, beta, ab)) +
                           (double)(n - 1 - i) * std::fabs(higham_entry(i, i + 1, alpha, beta, ab));
        best = std::max(best, row);
        left += std::fabs(higham_entry(i + 1, i, alpha, beta, ab));  // a_{i+1, i}: column i below row i+1
    }
    return best;
}

int omx_higham_matgen_f32(int64_t n, double alpha, double beta, float *A, int64_t lda, double *diag,
                          void *stream) {
    if (n <= 0 || lda < n || !A || !diag || !(std::isfinite(alpha) && std::isfinite(beta))) return OMX_ERR_ARG;
    const unsigned grid = (unsigned)std::min<int64_t>(n, 65535);
    hipLaunchKernelGGL(k_higham_fill, dim3(grid), dim3(256), 0, omx_stream(stream), n, alpha, beta, A, lda, diag);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

int omx_higham_gemv_f64(int64_t n, double alpha, double beta, double s, const double *x, double *y,
                        void *stream) {
    if (n <= 0 || !x || !y) return OMX_ERR_ARG;
    hipLaunchKernelGGL(k_higham_gemv, dim3((unsigned)omx_cdiv(n, 256)), dim3(256), 0, omx_stream(stream), n, alpha,
                       beta, s, x, y);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

}  // extern "C"
