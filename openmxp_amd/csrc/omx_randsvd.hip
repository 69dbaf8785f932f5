// omx_randsvd.hip -- the conditioned test-matrix family ("randsvd"):
//   A = U diag(sigma) V^T,  sigma_i = kappa^(-i/(n-1)),  i = 0 .. n-1,
// with U = H_0 H_1 ... H_{k-1} and V = H_k ... H_{2k-1} products of seeded
// Householder reflectors H_t = I - 2 w_t w_t^T.  The reference has no
// counterpart (its Higham generator, highammgen.hpp:4-30, is a different,
// unreachable family); the singular values are exact by construction, so
// kappa_2(A) = kappa in exact arithmetic.
//
// Structure used everywhere: every reflector adds one rank-1 term, so
//   A = diag(sigma) + P Q^T  with P, Q of n x R, R = 2k,
// which lets the device fill the FP32 matrix at HBM speed, gives the FP64
// diagonal and row sums |A| for the harness, and turns the refinement's
// residual product into O(n R) work.
//
// Reflector vectors: w_t[i] = the reference LCG's value at linear index
// n^2 + n + t n + i (after the matrix entries, index < n^2, and the
// right-hand side, index n^2 + i), normalised in FP64.
#include <algorithm>
#include <cmath>
#include <vector>

#include "omx_common.hpp"

namespace {

constexpr int kFillRows = 256;
constexpr int kLrBlocks = 512;
constexpr int kMaxRank = 16;

int fill_segments(int64_t n) {
    const int64_t nrb = omx_cdiv(n, kFillRows);
    int64_t s = omx_cdiv(16384, nrb);
    if (s > 64) s = 64;
    const int64_t smax = omx_cdiv(n, 256);
    if (s > smax) s = smax;
    if (s < 1) s = 1;
    return (int)s;
}

// A(i, j) = [i == j] sigma_i + sum_r P(i, r) Q(j, r) in FP64 (r in order,
// fused), stored as FP32 (RNE); diag_i = A(i, i) in FP64; the segment's
// sum_j |A(i, j)| goes to part[seg*n + i] (or rowabs with one segment).
template <int R>
__global__ void __launch_bounds__(kFillRows)
k_randsvd_fill(int64_t n, const double *__restrict__ sigma, const double *__restrict__ P,
               const double *__restrict__ Q, float *__restrict__ A, int64_t lda, double *__restrict__ diag,
               int64_t seglen, double *__restrict__ part) {
    const int64_t i = (int64_t)blockIdx.x * kFillRows + threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.y * seglen;
    const int64_t c1 = min(c0 + seglen, n);
    if (i >= n) return;
    double p[R > 0 ? R : 1];
#pragma unroll
    for (int r = 0; r < R; ++r) p[r] = P[i + r * n];
    const double si = sigma[i];
    double s = 0.0;
    for (int64_t j = c0; j < c1; ++j) {
        double v = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) v = __builtin_fma(p[r], Q[j + r * n], v);
        if (j == i) {
            v += si;
            diag[i] = v;
        }
        A[i + j * lda] = (float)v;
        s += fabs(v);
    }
    part[(int64_t)blockIdx.y * n + i] = s;
}

__global__ void k_rowsum_reduce(int64_t n, int nseg, const double *__restrict__ part, double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int g = 0; g < nseg; ++g) s += part[(int64_t)g * n + i];
    out[i] = s;
}

// t_r = sum_i Q(i, r) x_i: kLrBlocks partial sums per r (block b takes the
// rows b, b + kLrBlocks*256, ... in 256-thread strides), then a fixed-order
// final sum.
__global__ void __launch_bounds__(256)
k_lr_dot_partial(int64_t n, const double *__restrict__ Q, const double *__restrict__ x, double *__restrict__ part) {
    __shared__ double red[256];
    const int r = blockIdx.y;
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)kLrBlocks * 256)
        s = __builtin_fma(Q[i + r * n], x[i], s);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[r * kLrBlocks + blockIdx.x] = red[0];
}

__global__ void k_lr_dot_final(int R, const double *__restrict__ part, double *__restrict__ t) {
    const int r = threadIdx.x;
    if (r >= R) return;
    double s = 0.0;
    for (int b = 0; b < kLrBlocks; ++b) s += part[r * kLrBlocks + b];
    t[r] = s;
}

// y_i += alpha * (sigma_i x_i + sum_r P(i, r) t_r)
__global__ void k_lr_apply(int64_t n, int R, const double *__restrict__ sigma, const double *__restrict__ P,
                           const double *__restrict__ t, double alpha, const double *__restrict__ x,
                           double *__restrict__ y) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        double v = sigma[i] * x[i];
        for (int r = 0; r < R; ++r) v = __builtin_fma(P[i + r * n], t[r], v);
        y[i] = __builtin_fma(alpha, v, y[i]);
    }
}

template <int R>
void launch_fill(dim3 grid, hipStream_t st, int64_t n, const double *sigma, const double *P, const double *Q,
                 float *A, int64_t lda, double *diag, int64_t seglen, double *part) {
    hipLaunchKernelGGL(k_randsvd_fill<R>, grid, dim3(kFillRows), 0, st, n, sigma, P, Q, A, lda, diag, seglen,
                       part);
}

}  // namespace

extern "C" {

int omx_randsvd_factors_host(int64_t n, uint64_t seed, double kappa, int nrefl, double *sigma, double *P,
                             double *Q) {
    if (n < 1 || nrefl < 0 || 2 * nrefl > kMaxRank || !(kappa >= 1.0) || !std::isfinite(kappa) || !sigma ||
        (nrefl > 0 && (!P || !Q)))
        return OMX_ERR_ARG;
    for (int64_t i = 0; i < n; ++i)
        sigma[i] = (n == 1) ? 1.0 : std::pow(kappa, -(double)i / (double)(n - 1));
    const int k = nrefl;
    std::vector<double> w((size_t)n), c((size_t)2 * k + 1);
    const uint64_t x00 = omx_lcg_init(seed);
    auto make_w = [&](int t) {
        const uint64_t idx = (uint64_t)n * (uint64_t)n + (uint64_t)n + (uint64_t)t * (uint64_t)n;
        uint64_t s = omx_apply(omx_jump(idx), x00);
        const omx_map one = omx_jump(1);
        double nrm2 = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            w[i] = omx_lcg_value_host(s);
            nrm2 += w[i] * w[i];
            s = omx_apply(one, s);
        }
        const double inv = 1.0 / std::sqrt(nrm2);
        for (int64_t i = 0; i < n; ++i) w[i] *= inv;
    };
    int cols = 0;
    // A = H_0 ... H_{k-1} diag(sigma) H_{2k-1} ... H_k: right factors first,
    // innermost (H_{2k-1}) first: (S + P Q^T) H = S + P Q^T + p w^T with
    // p = -2 (sigma.w + P (Q^T w)).
    for (int t = 2 * k - 1; t >= k; --t) {
        make_w(t);
        for (int r = 0; r < cols; ++r) {
            double d = 0.0;
            for (int64_t i = 0; i < n; ++i) d += Q[i + r * n] * w[i];
            c[r] = d;
        }
        for (int64_t i = 0; i < n; ++i) {
            double v = sigma[i] * w[i];
            for (int r = 0; r < cols; ++r) v += P[i + r * n] * c[r];
            P[i + cols * n] = -2.0 * v;
            Q[i + cols * n] = w[i];
        }
        ++cols;
    }
    // left factors, innermost (H_{k-1}) first: H (S + P Q^T) = S + P Q^T +
    // w q^T with q = -2 (sigma.w + Q (P^T w)).
    for (int t = k - 1; t >= 0; --t) {
        make_w(t);
        for (int r = 0; r < cols; ++r) {
            double d = 0.0;
            for (int64_t i = 0; i < n; ++i) d += P[i + r * n] * w[i];
            c[r] = d;
        }
        for (int64_t j = 0; j < n; ++j) {
            double v = sigma[j] * w[j];
            for (int r = 0; r < cols; ++r) v += Q[j + r * n] * c[r];
            P[j + cols * n] = w[j];
            Q[j + cols * n] = -2.0 * v;
        }
        ++cols;
    }
    return OMX_OK;
}

int omx_randsvd_factors(int64_t n, uint64_t seed, double kappa, int nrefl, double *sigma, double *P, double *Q,
                        void *stream) {
    if (n < 1 || nrefl < 0 || 2 * nrefl > kMaxRank || !sigma || (nrefl > 0 && (!P || !Q))) return OMX_ERR_ARG;
    const int R = 2 * nrefl;
    std::vector<double> hs((size_t)n), hp((size_t)n * (R > 0 ? R : 1)), hq((size_t)n * (R > 0 ? R : 1));
    int rc = omx_randsvd_factors_host(n, seed, kappa, nrefl, hs.data(), hp.data(), hq.data());
    if (rc) return rc;
    hipStream_t st = omx_stream(stream);
    OMX_HIP_CHECK(hipMemcpyAsync(sigma, hs.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
    if (R > 0) {
        OMX_HIP_CHECK(hipMemcpyAsync(P, hp.data(), sizeof(double) * n * R, hipMemcpyHostToDevice, st));
        OMX_HIP_CHECK(hipMemcpyAsync(Q, hq.data(), sizeof(double) * n * R, hipMemcpyHostToDevice, st));
    }
    // the host vectors go out of scope: wait for the copies
    OMX_HIP_CHECK(hipStreamSynchronize(st));
    return OMX_OK;
}

size_t omx_randsvd_workspace_bytes(int64_t n) {
    if (n <= 0) return 0;
    return (size_t)fill_segments(n) * (size_t)n * sizeof(double);
}

int omx_randsvd_f32(int64_t n, int64_t rank, const double *sigma, const double *P, const double *Q, float *A,
                    int64_t lda, double *diag, double *rowabs, void *work, size_t work_bytes, void *stream) {
    if (n < 1 || rank < 0 || rank > kMaxRank || rank % 2 || !sigma || (rank > 0 && (!P || !Q)) || !A ||
        lda < n || !diag || !rowabs)
        return OMX_ERR_ARG;
    const int nseg = fill_segments(n);
    const size_t need = omx_randsvd_workspace_bytes(n);
    if (!work || work_bytes < need) return OMX_ERR_WORKSPACE;
    hipStream_t st = omx_stream(stream);
    const int64_t seglen = omx_cdiv(n, nseg);
    dim3 grid((unsigned)omx_cdiv(n, kFillRows), (unsigned)nseg);
    double *part = static_cast<double *>(work);
    switch (rank) {
        case 0: launch_fill<0>(grid, st, n, sigma, P, Q, A, lda, diag, seglen, part); break;
        case 2: launch_fill<2>(grid, st, n, sigma, P, Q, A, lda, diag, seglen, part); break;
        case 4: launch_fill<4>(grid, st, n, sigma, P, Q, A, lda, diag, seglen, part); break;
        case 6: launch_fill<6>(grid, st, n, sigma, P, Q, A, lda, diag, seglen, part); break;
        case 8: launch_fill<8>(grid, st, n, sigma, P, Q, A, lda, diag, seglen, part); break;
        case 10: launch_fill<10>(grid, st, n, sigma, P, Q, A, lda, diag, seglen, part); break;
        case 12: launch_fill<12>(grid, st, n, sigma, P, Q, A, lda, diag, seglen, part); break;
        case 14: launch_fill<14>(grid, st, n, sigma, P, Q, A, lda, diag, seglen, part); break;
        default: launch_fill<16>(grid, st, n, sigma, P, Q, A, lda, diag, seglen, part); break;
    }
    OMX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_rowsum_reduce, dim3((unsigned)omx_cdiv(n, 256)), dim3(256), 0, st, n, nseg, part, rowabs);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

size_t omx_lowrank_gemv_workspace_bytes(int64_t rank) {
    return rank > 0 ? (size_t)(kLrBlocks + 1) * (size_t)rank * sizeof(double) : 0;
}

int omx_lowrank_gemv_f64(int64_t n, int64_t rank, const double *sigma, const double *P, const double *Q,
                         double alpha, const double *x, double *y, void *work, size_t work_bytes, void *stream) {
    if (n < 1 || rank < 0 || rank > kMaxRank || !sigma || (rank > 0 && (!P || !Q)) || !x || !y)
        return OMX_ERR_ARG;
    if (rank > 0 && (!work || work_bytes < omx_lowrank_gemv_workspace_bytes(rank))) return OMX_ERR_WORKSPACE;
    hipStream_t st = omx_stream(stream);
    double *part = static_cast<double *>(work);
    double *t = rank > 0 ? part + kLrBlocks * rank : nullptr;
    if (rank > 0) {
        hipLaunchKernelGGL(k_lr_dot_partial, dim3(kLrBlocks, (unsigned)rank), dim3(256), 0, st, n, Q, x, part);
        OMX_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_lr_dot_final, dim3(1), dim3(64), 0, st, (int)rank, part, t);
        OMX_LAUNCH_CHECK();
    }
    const int64_t g = std::min<int64_t>(omx_cdiv(n, 256), 65536);
    hipLaunchKernelGGL(k_lr_apply, dim3((unsigned)g), dim3(256), 0, st, n, (int)rank, sigma, P, t, alpha, x, y);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

}  // extern "C"
