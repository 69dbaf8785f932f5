// omx_ir.hip -- FP64 kernels of the iterative refinement: the on-the-fly
// residual GEMV (A regenerated from the LCG), infinity norms, vector ops and
// the blocked triangular solves with the FP32 LU factors.
#include <float.h>

#include "omx_common.hpp"

namespace {

constexpr int kGemvRows = 256;     // rows per workgroup (one per thread)
constexpr int kGemvTargetWG = 16384;
constexpr int kGemvMaxSeg = 64;

// Number of column segments of the residual GEMV for order n: enough
// workgroups to fill the 256 CUs several times, each segment >= 256 columns.
int gemv_segments(int64_t n) {
    const int64_t nrb = omx_cdiv(n, kGemvRows);
    int64_t s = omx_cdiv(kGemvTargetWG, nrb);
    if (s > kGemvMaxSeg) s = kGemvMaxSeg;
    const int64_t smax = omx_cdiv(n, 256);
    if (s > smax) s = smax;
    if (s < 1) s = 1;
    return (int)s;
}

// Row i of the workgroup's 256-row block, columns [c0, c1): walk along the
// row with the map a^n (one affine step per column).  The diagonal column
// takes diag[i] instead of the LCG value (otf_gemv_diag_d,
// panel_gemv.hpp:13-47).  part[seg*n + i] = the segment's FP64 sum, or, with
// a single segment, y_i += alpha * sum directly.
__global__ void __launch_bounds__(kGemvRows)
k_otf_gemv(int64_t n, uint64_t x00, omx_map jn, const double *__restrict__ diag,
           const double *__restrict__ x, int64_t seglen, double *__restrict__ part,
           double *__restrict__ y, double alpha) {
    const int64_t i = (int64_t)blockIdx.x * kGemvRows + threadIdx.x;
    const int64_t seg = blockIdx.y;
    const int64_t c0 = seg * seglen;
    int64_t c1 = c0 + seglen;
    if (c1 > n) c1 = n;
    if (i >= n) return;
    uint64_t s = omx_apply(omx_jump((uint64_t)i + (uint64_t)n * (uint64_t)c0), x00);
    double acc = 0.0;
    // diagonal columns of this row block
    const int64_t d0 = (int64_t)blockIdx.x * kGemvRows;
    const int64_t d1 = d0 + kGemvRows;
    const int64_t lo = c0 > d0 ? c0 : (d0 < c1 ? d0 : c1);
    const int64_t hi = c1 < d1 ? c1 : d1;
    int64_t j = c0;
    for (; j < lo; ++j) {
        acc = __builtin_fma(omx_lcg_value(s), x[j], acc);
        s = omx_apply(jn, s);
    }
    const double di = diag[i];
    for (; j < hi; ++j) {
        const double a = (j == i) ? di : omx_lcg_value(s);
        acc = __builtin_fma(a, x[j], acc);
        s = omx_apply(jn, s);
    }
    for (; j < c1; ++j) {
        acc = __builtin_fma(omx_lcg_value(s), x[j], acc);
        s = omx_apply(jn, s);
    }
    if (part)
        part[seg * n + i] = acc;
    else
        y[i] = y[i] + alpha * acc;
}

__global__ void k_gemv_reduce(int64_t n, int nseg, const double *__restrict__ part,
                              double *__restrict__ y, double alpha) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = part[i];
    for (int q = 1; q < nseg; ++q) s += part[(int64_t)q * n + i];
    y[i] = y[i] + alpha * s;
}

__global__ void __launch_bounds__(1024)
k_norm_inf(int64_t n, const double *__restrict__ x, double *__restrict__ out) {
    __shared__ double sh[1024];
    double m = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 1024) m = fmax(m, fabs(x[i]));
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

__global__ void k_copy(int64_t n, const double *__restrict__ x, double *__restrict__ y) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] = x[i];
}

__global__ void k_div(int64_t n, const double *__restrict__ b, const double *__restrict__ d,
                      double *__restrict__ x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = b[i] / d[i];
}

__global__ void k_axpy(int64_t n, double alpha, const double *__restrict__ x,
                       double *__restrict__ y) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] = y[i] + alpha * x[i];
}

__global__ void k_zero(int64_t n, double *__restrict__ y) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Blocked triangular solves.  The diagonal b x b block is solved by one
// workgroup of 1024 threads in sub-blocks of 64: wave 0 solves the 64 x 64
// triangle with lane-to-lane broadcasts, then all threads update the rows
// of the block below (lower) or above (upper) the sub-block.  The vector
// lives in LDS (b <= kTrsvMaxB).
// ---------------------------------------------------------------------------
constexpr int kTrsvThreads = 1024;
constexpr int kTrsvSub = 64;
constexpr int kTrsvMaxB = 8192;

__global__ void __launch_bounds__(kTrsvThreads)
k_trsv_diag_lower(int64_t b, const float *__restrict__ A, int64_t lda,
                  double *__restrict__ x, const double *__restrict__ acc) {
    extern __shared__ double xs[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    for (int64_t i = tid; i < b; i += kTrsvThreads) xs[i] = x[i] - acc[i];
    __syncthreads();
    for (int64_t sb = 0; sb < b; sb += kTrsvSub) {
        if (wave == 0) {
            const int64_t row = sb + lane;
            float a[kTrsvSub];
#pragma unroll
            for (int l = 0; l < kTrsvSub; ++l) a[l] = (lane > l) ? A[row + (sb + l) * lda] : 0.0f;
            double v = xs[row];
#pragma unroll
            for (int l = 0; l < kTrsvSub; ++l) {
                const double t = __shfl(v, l, 64);
                if (lane > l) v -= t * (double)a[l];
            }
            xs[row] = v;
        }
        __syncthreads();
        for (int64_t i = sb + kTrsvSub + tid; i < b; i += kTrsvThreads) {
            double s = 0.0;
#pragma unroll 16
            for (int l = 0; l < kTrsvSub; ++l) s += (double)A[i + (sb + l) * lda] * xs[sb + l];
            xs[i] -= s;
        }
        __syncthreads();
    }
    for (int64_t i = tid; i < b; i += kTrsvThreads) x[i] = xs[i];
}

__global__ void __launch_bounds__(kTrsvThreads)
k_trsv_diag_upper(int64_t b, const float *__restrict__ A, int64_t lda,
                  double *__restrict__ x, const double *__restrict__ acc) {
    extern __shared__ double xs[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    for (int64_t i = tid; i < b; i += kTrsvThreads) xs[i] = x[i] - acc[i];
    __syncthreads();
    for (int64_t sb = b - kTrsvSub; sb >= 0; sb -= kTrsvSub) {
        if (wave == 0) {
            const int64_t row = sb + lane;
            float a[kTrsvSub];
#pragma unroll
            for (int l = 0; l < kTrsvSub; ++l) a[l] = (lane <= l) ? A[row + (sb + l) * lda] : 0.0f;
            double v = xs[row];
#pragma unroll
            for (int l = kTrsvSub - 1; l >= 0; --l) {
                if (lane == l) v = v / (double)a[l];
                const double t = __shfl(v, l, 64);
                if (lane < l) v -= t * (double)a[l];
            }
            xs[row] = v;
        }
        __syncthreads();
        for (int64_t i = tid; i < sb; i += kTrsvThreads) {
            double s = 0.0;
#pragma unroll 16
            for (int l = 0; l < kTrsvSub; ++l) s += (double)A[i + (sb + l) * lda] * xs[sb + l];
            xs[i] -= s;
        }
        __syncthreads();
    }
    for (int64_t i = tid; i < b; i += kTrsvThreads) x[i] = xs[i];
}

// acc_i += sum_l A(i, l) * x_l for rows i < m, l < b (FP32 matrix, FP64
// vector and sums); one thread per row, columns in order.
__global__ void __launch_bounds__(256)
k_trsv_gemv(int64_t m, int64_t b, const float *__restrict__ A, int64_t lda,
            const double *__restrict__ xv, double *__restrict__ acc) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    double s = 0.0;
    const float *p = A + i;
#pragma unroll 8
    for (int64_t l = 0; l < b; ++l) s += (double)p[l * lda] * xv[l];
    acc[i] += s;
}

int grid_for(int64_t n, int threads) {
    int64_t g = omx_cdiv(n, threads);
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (int)g;
}

}  // namespace

extern "C" {

size_t omx_otf_gemv_workspace_bytes(int64_t n) {
    if (n <= 0) return 0;
    const int s = gemv_segments(n);
    return s > 1 ? (size_t)s * (size_t)n * sizeof(double) : 0;
}

int omx_otf_gemv_f64(int64_t n, uint64_t seed, const double *diag, double alpha,
                     const double *x, double *y, void *work, size_t work_bytes,
                     void *stream) {
    if (n <= 0 || !diag || !x || !y) return OMX_ERR_ARG;
    const int nseg = gemv_segments(n);
    const size_t need = omx_otf_gemv_workspace_bytes(n);
    if (need > 0 && (!work || work_bytes < need)) return OMX_ERR_WORKSPACE;
    hipStream_t st = omx_stream(stream);
    const int64_t seglen = omx_cdiv(n, nseg);
    const int64_t nrb = omx_cdiv(n, kGemvRows);
    double *part = nseg > 1 ? static_cast<double *>(work) : nullptr;
    hipLaunchKernelGGL(k_otf_gemv, dim3((unsigned)nrb, (unsigned)nseg), dim3(kGemvRows), 0, st, n,
                       omx_lcg_init(seed), omx_jump((uint64_t)n), diag, x, seglen, part, y, alpha);
    OMX_LAUNCH_CHECK();
    if (nseg > 1) {
        hipLaunchKernelGGL(k_gemv_reduce, dim3((unsigned)omx_cdiv(n, 256)), dim3(256), 0, st, n,
                           nseg, part, y, alpha);
        OMX_LAUNCH_CHECK();
    }
    return OMX_OK;
}

int omx_norm_inf_f64(int64_t n, const double *x, double *result, void *stream) {
    if (n <= 0 || !x || !result) return OMX_ERR_ARG;
    hipLaunchKernelGGL(k_norm_inf, dim3(1), dim3(1024), 0, omx_stream(stream), n, x, result);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

int omx_copy_f64(int64_t n, const double *x, double *y, void *stream) {
    if (n <= 0 || !x || !y) return OMX_ERR_ARG;
    hipLaunchKernelGGL(k_copy, dim3(grid_for(n, 256)), dim3(256), 0, omx_stream(stream), n, x, y);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

int omx_div_f64(int64_t n, const double *b, const double *d, double *x, void *stream) {
    if (n <= 0 || !b || !d || !x) return OMX_ERR_ARG;
    hipLaunchKernelGGL(k_div, dim3(grid_for(n, 256)), dim3(256), 0, omx_stream(stream), n, b, d, x);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

int omx_axpy_f64(int64_t n, double alpha, const double *x, double *y, void *stream) {
    if (n <= 0 || !x || !y) return OMX_ERR_ARG;
    hipLaunchKernelGGL(k_axpy, dim3(grid_for(n, 256)), dim3(256), 0, omx_stream(stream), n, alpha,
                       x, y);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

size_t omx_trsv_workspace_bytes(int64_t n) { return n > 0 ? (size_t)n * sizeof(double) : 0; }

static int trsv_args(int64_t n, int64_t b, const float *A, int64_t lda, double *x, void *work,
                     size_t work_bytes) {
    if (n <= 0 || b <= 0 || n % b || b % kTrsvSub || b > kTrsvMaxB || lda < n || !A || !x)
        return OMX_ERR_ARG;
    if (!work || work_bytes < omx_trsv_workspace_bytes(n)) return OMX_ERR_WORKSPACE;
    return OMX_OK;
}

int omx_trsv_lower_unit_f32f64(int64_t n, int64_t b, const float *A, int64_t lda, double *x,
                               void *work, size_t work_bytes, void *stream) {
    int rc = trsv_args(n, b, A, lda, x, work, work_bytes);
    if (rc) return rc;
    hipStream_t st = omx_stream(stream);
    double *acc = static_cast<double *>(work);
    hipLaunchKernelGGL(k_zero, dim3(grid_for(n, 256)), dim3(256), 0, st, n, acc);
    OMX_LAUNCH_CHECK();
    const size_t lds = (size_t)b * sizeof(double);
    for (int64_t j0 = 0; j0 < n; j0 += b) {
        hipLaunchKernelGGL(k_trsv_diag_lower, dim3(1), dim3(kTrsvThreads), lds, st, b,
                           A + j0 + j0 * lda, lda, x + j0, acc + j0);
        OMX_LAUNCH_CHECK();
        const int64_t rest = n - j0 - b;
        if (rest > 0) {
            hipLaunchKernelGGL(k_trsv_gemv, dim3((unsigned)omx_cdiv(rest, 256)), dim3(256), 0, st,
                               rest, b, A + (j0 + b) + j0 * lda, lda, x + j0, acc + j0 + b);
            OMX_LAUNCH_CHECK();
        }
    }
    return OMX_OK;
}

int omx_trsv_upper_f32f64(int64_t n, int64_t b, const float *A, int64_t lda, double *x,
                          void *work, size_t work_bytes, void *stream) {
    int rc = trsv_args(n, b, A, lda, x, work, work_bytes);
    if (rc) return rc;
    hipStream_t st = omx_stream(stream);
    double *acc = static_cast<double *>(work);
    hipLaunchKernelGGL(k_zero, dim3(grid_for(n, 256)), dim3(256), 0, st, n, acc);
    OMX_LAUNCH_CHECK();
    const size_t lds = (size_t)b * sizeof(double);
    for (int64_t j0 = n - b; j0 >= 0; j0 -= b) {
        hipLaunchKernelGGL(k_trsv_diag_upper, dim3(1), dim3(kTrsvThreads), lds, st, b,
                           A + j0 + j0 * lda, lda, x + j0, acc + j0);
        OMX_LAUNCH_CHECK();
        if (j0 > 0) {
            hipLaunchKernelGGL(k_trsv_gemv, dim3((unsigned)omx_cdiv(j0, 256)), dim3(256), 0, st,
                               j0, b, A + j0 * lda, lda, x + j0, acc);
            OMX_LAUNCH_CHECK();
        }
    }
    return OMX_OK;
}

double omx_harness(double normr, double norma, double normx, double normb, int64_t n) {
    return normr / (norma * normx + normb) * 1.0 / ((double)n * DBL_EPSILON / 2);
}

}  // extern "C"
