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
// Triangular solves with the FP32 factors, FP64 vector and arithmetic.
//
// The matrix is walked in blocks of kTB = 256 rows/columns.  One launch per
// block column J (solved values x_J final in x):
//   workgroup 0   rows of the next diagonal block D: r_D -= M(D, J) x_J, then
//                 the diagonal block is solved in LDS (4 sub-blocks of 64,
//                 each a lane-to-lane chain in one wave, followed by the
//                 update of the later sub-blocks' rows), x_D written back;
//   workgroup w   the remaining rows beyond D, 256 per workgroup:
//                 r -= M(rows, J) x_J.
// So the diagonal chain of block D runs concurrently with the bulk update by
// block J, and the solve reads each FP32 factor entry exactly once (HBM-bound:
// 2 n^2 bytes per triangle).  Summation order per row and block column: the
// four waves each sum 64 columns in order, the partial sums are added in
// wave order, the total is subtracted from r.
// ---------------------------------------------------------------------------
constexpr int kTB = 256;
constexpr int kSB = 64;

// Lane l's value of v in every lane (l uniform): two readlanes.
__device__ __forceinline__ double bcast_lane(double v, int l) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// sum over the block's columns of M(row, c) x_c for 256 rows from r0 (rows
// >= rmax ignored) and ncols <= 256 columns from c0; xv (LDS) holds x of the
// columns.  Wave w takes columns [64 w, 64 w + 64), lane l rows 4l .. 4l+3
// (one 16-byte load per column when kVec).  Returns the row's total in
// thread t = row - r0 (after the LDS reduction).
template <bool kVec>
__device__ __forceinline__ double block_rows_dot(const float *__restrict__ A, int64_t lda, int64_t r0,
                                                 int64_t rmax, int64_t c0, int ncols, const double *xv,
                                                 double (*part)[kTB]) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int cb = wave * kSB;
    const int ce = min(ncols, cb + kSB);
    const int64_t row = r0 + 4 * lane;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    if (kVec && row + 3 < rmax) {
        const float *p = A + row + (c0 + cb) * lda;
        int c = cb;
        for (; c + 8 <= ce; c += 8) {
            float4 v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = *reinterpret_cast<const float4 *>(p + (int64_t)(c - cb + u) * lda);
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const double xc = xv[c + u];
                s0 = __builtin_fma((double)v[u].x, xc, s0);
                s1 = __builtin_fma((double)v[u].y, xc, s1);
                s2 = __builtin_fma((double)v[u].z, xc, s2);
                s3 = __builtin_fma((double)v[u].w, xc, s3);
            }
        }
        for (; c < ce; ++c) {
            const float4 v = *reinterpret_cast<const float4 *>(p + (int64_t)(c - cb) * lda);
            const double xc = xv[c];
            s0 = __builtin_fma((double)v.x, xc, s0);
            s1 = __builtin_fma((double)v.y, xc, s1);
            s2 = __builtin_fma((double)v.z, xc, s2);
            s3 = __builtin_fma((double)v.w, xc, s3);
        }
    } else {
        for (int c = cb; c < ce; ++c) {
            const float *q = A + row + (c0 + c) * lda;
            const double xc = xv[c];
            if (row < rmax) s0 = __builtin_fma((double)q[0], xc, s0);
            if (row + 1 < rmax) s1 = __builtin_fma((double)q[1], xc, s1);
            if (row + 2 < rmax) s2 = __builtin_fma((double)q[2], xc, s2);
            if (row + 3 < rmax) s3 = __builtin_fma((double)q[3], xc, s3);
        }
    }
    part[wave][4 * lane] = s0;
    part[wave][4 * lane + 1] = s1;
    part[wave][4 * lane + 2] = s2;
    part[wave][4 * lane + 3] = s3;
    __syncthreads();
    const double t = ((part[0][tid] + part[1][tid]) + part[2][tid]) + part[3][tid];
    __syncthreads();
    return t;
}

// Lane's row `row` of the diagonal block at d0 (order dim), the 64 columns
// from c0: row and column indices are clamped into the block, so every load
// is unconditional and in bounds (all 64 can be in flight at once); entries
// outside the block are masked where they are used.
__device__ __forceinline__ void load_row64(const float *__restrict__ A, int64_t lda, int64_t d0, int dim,
                                           int row, int c0, float (&a)[kSB]) {
    const float *p = A + d0 + min(row, dim - 1) + d0 * lda;
#pragma unroll
    for (int c = 0; c < kSB; ++c) a[c] = p[(int64_t)min(c0 + c, dim - 1) * lda];
}

// x_D <- L_DD^-1 r_D for the unit-lower diagonal block at d0 of order dim
// (<= 256); r_D in xs (LDS).  Wave w owns rows [64 w, 64 w + 64) (lane l:
// row 64 w + l) and holds its diagonal sub-block's rows in registers from
// the start.  Sub-block sb: wave sb runs the 64-step chain (x_c read from
// lane c), then every later wave subtracts its rows' product with x_sb,
// using the off-diagonal sub-block it prefetched during the previous stage.
__device__ __forceinline__ void diag_solve_lower(const float *__restrict__ A, int64_t lda, int64_t d0, int dim,
                                                 double *xs) {
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nsb = (dim + kSB - 1) / kSB;
    const int i = wave * kSB + lane;
    float a[kSB], b[kSB];
    if (wave < nsb) {
        load_row64(A, lda, d0, dim, i, wave * kSB, a);
        if (wave > 0) load_row64(A, lda, d0, dim, i, 0, b);
    }
#pragma unroll
    for (int sb = 0; sb < kTB / kSB; ++sb) {
        if (sb >= nsb) break;
        const int c0 = sb * kSB;
        if (wave == sb) {
            double v = xs[i];
#pragma unroll
            for (int c = 0; c < kSB; ++c) {
                const double m = (lane > c && i < dim) ? (double)a[c] : 0.0;
                v = __builtin_fma(-bcast_lane(v, c), m, v);
            }
            if (i < dim) xs[i] = v;
        }
        __syncthreads();
        if (wave > sb && wave < nsb) {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < kSB; ++c) s = __builtin_fma(c0 + c < dim ? (double)b[c] : 0.0, xs[c0 + c], s);
            if (i < dim) xs[i] -= s;
            if (sb + 1 < wave) load_row64(A, lda, d0, dim, i, c0 + kSB, b);  // next stage's sub-block
        }
        __syncthreads();
    }
}

// x_D <- U_DD^-1 r_D for the upper non-unit diagonal block (sub-blocks from
// the last to the first; registers and prefetch as in the lower solve).  The
// quotient r / u is formed as q0 = r * y, q = q0 + (r - u q0) y with
// y = 1/u: the correctly rounded quotient (Markstein's correction with a
// correctly rounded reciprocal).
__device__ __forceinline__ void diag_solve_upper(const float *__restrict__ A, int64_t lda, int64_t d0, int dim,
                                                 double *xs) {
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nsb = (dim + kSB - 1) / kSB;
    const int i = wave * kSB + lane;
    float a[kSB], b[kSB];
    float ud = 1.0f;
    if (wave < nsb) {
        load_row64(A, lda, d0, dim, i, wave * kSB, a);
        ud = A[d0 + min(i, dim - 1) + (d0 + min(i, dim - 1)) * lda];
        if (wave < nsb - 1) load_row64(A, lda, d0, dim, i, (nsb - 1) * kSB, b);
    }
#pragma unroll
    for (int t = 0; t < kTB / kSB; ++t) {
        const int sb = nsb - 1 - t;
        if (sb < 0) break;
        const int c0 = sb * kSB;
        if (wave == sb) {
            const double u = (i < dim) ? (double)ud : 1.0;
            const double y = 1.0 / u;
            auto quot = [&](double r) {
                const double q0 = r * y;
                return __builtin_fma(__builtin_fma(-u, q0, r), y, q0);
            };
            // Step c: every lane forms its quotient, lane c's is x_c; rows
            // above subtract x_c u(i, c).  Lane c's own v is untouched from
            // then on, so its final quotient below equals x_c.
            double v = (i < dim) ? xs[i] : 0.0;
#pragma unroll
            for (int c = kSB - 1; c >= 0; --c) {
                const double m = (lane < c && c0 + c < dim) ? (double)a[c] : 0.0;
                v = __builtin_fma(-bcast_lane(quot(v), c), m, v);
            }
            if (i < dim) xs[i] = quot(v);
        }
        __syncthreads();
        if (wave < sb) {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < kSB; ++c) s = __builtin_fma(c0 + c < dim ? (double)b[c] : 0.0, xs[c0 + c], s);
            xs[i] -= s;
            if (sb - 1 > wave) load_row64(A, lda, d0, dim, i, c0 - kSB, b);  // next stage's sub-block
        }
        __syncthreads();
    }
}

// One block column of the forward solve.  j0 = first column of the solved
// block J (j0 < 0: nothing solved yet, only the first diagonal block).
template <bool kVec>
__global__ void __launch_bounds__(256)
k_trsv_lower_step(int64_t n, int64_t j0, const float *__restrict__ A, int64_t lda, double *__restrict__ x) {
    __shared__ double xv[kTB];
    __shared__ double xs[kTB];
    __shared__ double part[4][kTB];
    const int tid = threadIdx.x;
    const int64_t jb = j0 < 0 ? 0 : j0;
    const int ncols = j0 < 0 ? 0 : (int)min((int64_t)kTB, n - j0);
    const int64_t d0 = j0 < 0 ? 0 : j0 + kTB;
    if (ncols > 0) xv[tid] = tid < ncols ? x[jb + tid] : 0.0;
    __syncthreads();
    if (blockIdx.x == 0) {
        const int dim = (int)min((int64_t)kTB, n - d0);
        double r = (tid < dim) ? x[d0 + tid] : 0.0;
        if (ncols > 0) r -= block_rows_dot<kVec>(A, lda, d0, d0 + dim, jb, ncols, xv, part);
        xs[tid] = r;
        __syncthreads();
        diag_solve_lower(A, lda, d0, dim, xs);
        if (tid < dim) x[d0 + tid] = xs[tid];
        return;
    }
    const int64_t r0 = d0 + (int64_t)blockIdx.x * kTB;
    const double t = block_rows_dot<kVec>(A, lda, r0, n, jb, ncols, xv, part);
    if (r0 + tid < n) x[r0 + tid] -= t;
}

// One block column of the backward solve: J = [j0, j0 + ncols) solved
// (j0 >= n: nothing solved yet, only the last diagonal block, which may be
// shorter than kTB); D = the block before J; rows [0, d0) updated.
template <bool kVec>
__global__ void __launch_bounds__(256)
k_trsv_upper_step(int64_t n, int64_t j0, const float *__restrict__ A, int64_t lda, double *__restrict__ x) {
    __shared__ double xv[kTB];
    __shared__ double xs[kTB];
    __shared__ double part[4][kTB];
    const int tid = threadIdx.x;
    const int64_t last0 = ((n - 1) / kTB) * kTB;  // first row of the last block
    const bool none = j0 >= n;
    const int ncols = none ? 0 : (int)min((int64_t)kTB, n - j0);
    const int64_t d0 = none ? last0 : j0 - kTB;
    const int dim = (int)min((int64_t)kTB, n - d0);
    if (ncols > 0) xv[tid] = tid < ncols ? x[j0 + tid] : 0.0;
    __syncthreads();
    if (blockIdx.x == 0) {
        double r = (tid < dim) ? x[d0 + tid] : 0.0;
        if (ncols > 0) r -= block_rows_dot<kVec>(A, lda, d0, d0 + dim, j0, ncols, xv, part);
        xs[tid] = r;
        __syncthreads();
        diag_solve_upper(A, lda, d0, dim, xs);
        if (tid < dim) x[d0 + tid] = xs[tid];
        return;
    }
    const int64_t r0 = (int64_t)(blockIdx.x - 1) * kTB;
    const double t = block_rows_dot<kVec>(A, lda, r0, d0, j0, ncols, xv, part);
    if (r0 + tid < d0) x[r0 + tid] -= t;
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

size_t omx_trsv_workspace_bytes(int64_t n) {
    (void)n;
    return 0;
}

static int trsv_args(int64_t n, int64_t b, const float *A, int64_t lda, double *x) {
    if (n <= 0 || b <= 0 || n % b || lda < n || !A || !x) return OMX_ERR_ARG;
    return OMX_OK;
}

static bool trsv_vec(const float *A, int64_t lda) {
    return (reinterpret_cast<uintptr_t>(A) & 15) == 0 && (lda & 3) == 0;
}

int omx_trsv_lower_unit_f32f64(int64_t n, int64_t b, const float *A, int64_t lda, double *x,
                               void *work, size_t work_bytes, void *stream) {
    (void)work;
    (void)work_bytes;
    int rc = trsv_args(n, b, A, lda, x);
    if (rc) return rc;
    hipStream_t st = omx_stream(stream);
    const bool vec = trsv_vec(A, lda);
    for (int64_t j0 = -kTB; j0 + kTB < n; j0 += kTB) {
        const int64_t d0 = j0 + kTB;
        const int64_t beyond = n - d0 - kTB;
        const unsigned grid = 1u + (unsigned)(beyond > 0 && j0 >= 0 ? omx_cdiv(beyond, kTB) : 0);
        if (vec)
            hipLaunchKernelGGL(k_trsv_lower_step<true>, dim3(grid), dim3(256), 0, st, n, j0, A, lda, x);
        else
            hipLaunchKernelGGL(k_trsv_lower_step<false>, dim3(grid), dim3(256), 0, st, n, j0, A, lda, x);
        OMX_LAUNCH_CHECK();
    }
    return OMX_OK;
}

int omx_trsv_upper_f32f64(int64_t n, int64_t b, const float *A, int64_t lda, double *x,
                          void *work, size_t work_bytes, void *stream) {
    (void)work;
    (void)work_bytes;
    int rc = trsv_args(n, b, A, lda, x);
    if (rc) return rc;
    hipStream_t st = omx_stream(stream);
    const bool vec = trsv_vec(A, lda);
    const int64_t last0 = ((n - 1) / kTB) * kTB;
    // j0 = n: the last block alone; then J = each block from the last down
    for (int64_t j0 = n; j0 > 0; j0 = (j0 == n) ? last0 : j0 - kTB) {
        const int64_t d0 = (j0 == n) ? last0 : j0 - kTB;
        if (d0 < 0) break;
        const unsigned grid = 1u + (unsigned)(j0 < n ? omx_cdiv(d0, kTB) : 0);
        if (vec)
            hipLaunchKernelGGL(k_trsv_upper_step<true>, dim3(grid), dim3(256), 0, st, n, j0, A, lda, x);
        else
            hipLaunchKernelGGL(k_trsv_upper_step<false>, dim3(grid), dim3(256), 0, st, n, j0, A, lda, x);
        OMX_LAUNCH_CHECK();
    }
    return OMX_OK;
}

double omx_harness(double normr, double norma, double normx, double normb, int64_t n) {
    return normr / (norma * normx + normb) * 1.0 / ((double)n * DBL_EPSILON / 2);
}

}  // extern "C"
