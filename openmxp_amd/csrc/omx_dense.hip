// omx_dense.hip -- FP32 dense kernels on the LU critical path: the no-pivot
// LU of the b x b diagonal block (panel_trf), the triangular solves with its
// factors (panel_trsm and the -alt inverses), and the FP32 GEMM that updates
// inside them.
//
// All are blocked by 64: a 64-wide diagonal piece is handled by one small
// kernel (rows or columns in parallel, the 64 x 64 triangle staged in LDS),
// the rest of the update is an FP32 GEMM.  Pivots are applied as
// multiplication by the reciprocal, as sgetrf_nopiv_naive does
// (sgetrf_nopiv.cpp:27-61) and as reference BLAS strsm does for a non-unit
// diagonal.
#include "omx_common.hpp"

namespace {

constexpr int kNB = 64;

// ---------------------------------------------------------------------------
// FP32 GEMM, C = beta*C + alpha * A * B (all column-major, no transposes).
// 64 x 64 tile per workgroup of 256 threads, 4 x 4 results per thread,
// k advanced 16 at a time through LDS.  k must be a multiple of 16.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_sgemm_nn(int64_t M, int64_t N, int64_t K, float alpha, const float *__restrict__ A, int64_t lda,
           const float *__restrict__ B, int64_t ldb, float beta, float *__restrict__ C, int64_t ldc) {
    __shared__ float As[16][64 + 4];
    __shared__ float Bs[16][64 + 4];
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int64_t i0 = (int64_t)blockIdx.x * 64, j0 = (int64_t)blockIdx.y * 64;
    float acc[4][4] = {};
    for (int64_t k0 = 0; k0 < K; k0 += 16) {
        // A tile: 64 rows x 16 columns; thread -> (row = tid % 64, 4 columns)
        {
            const int r = tid & 63, c4 = (tid >> 6) * 4;
            const int64_t gi = i0 + r;
#pragma unroll
            for (int c = 0; c < 4; ++c) As[c4 + c][r] = (gi < M) ? A[gi + (k0 + c4 + c) * lda] : 0.0f;
        }
        // B tile: 16 rows x 64 columns; thread -> (column = tid / 4, 4 rows)
        {
            const int cc = tid >> 2, r4 = (tid & 3) * 4;
            const int64_t gj = j0 + cc;
#pragma unroll
            for (int r = 0; r < 4; ++r) Bs[r4 + r][cc] = (gj < N) ? B[(k0 + r4 + r) + gj * ldb] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a[q] = As[kk][tx + 16 * q];
                b[q] = Bs[kk][ty + 16 * q];
            }
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[p][q] = __builtin_fmaf(a[p], b[q], acc[p][q]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const int64_t gi = i0 + tx + 16 * p;
        if (gi >= M) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t gj = j0 + ty + 16 * q;
            if (gj >= N) continue;
            float *c = C + gi + gj * ldc;
            *c = (beta == 0.0f) ? alpha * acc[p][q] : beta * *c + alpha * acc[p][q];
        }
    }
}

// ---------------------------------------------------------------------------
// 64 x 64 diagonal piece of the LU: unblocked right-looking LU in LDS.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_getrf_diag64(float *__restrict__ A, int64_t lda, int64_t row0, int *__restrict__ info) {
    __shared__ float a[kNB][kNB + 1];
    const int tid = threadIdx.x;
    for (int e = tid; e < kNB * kNB; e += 256) a[e & 63][e >> 6] = A[(e & 63) + (int64_t)(e >> 6) * lda];
    __syncthreads();
    for (int k = 0; k < kNB; ++k) {
        const float akk = a[k][k];
        if (tid == 0 && akk == 0.0f && info) atomicMin(info, (int)(row0 + k + 1));
        const float inv = 1.0f / akk;
        if (tid > k && tid < kNB) a[tid][k] *= inv;
        __syncthreads();
        for (int e = tid; e < kNB * kNB; e += 256) {
            const int i = e & 63, j = e >> 6;
            if (i > k && j > k) a[i][j] -= a[i][k] * a[k][j];
        }
        __syncthreads();
    }
    for (int e = tid; e < kNB * kNB; e += 256) A[(e & 63) + (int64_t)(e >> 6) * lda] = a[e & 63][e >> 6];
}

// Rows: X(i, :) <- X(i, :) * U^-1 for a 64 x 64 upper non-unit U (RUNN).
// Column j of the row is b_j minus earlier columns, times 1/U_jj.
__global__ void __launch_bounds__(256)
k_trsm_rows_runn64(int64_t m, const float *__restrict__ U, int64_t ldu, float *__restrict__ X,
                   int64_t ldx) {
    __shared__ float u[kNB][kNB + 1];
    const int tid = threadIdx.x;
    for (int e = tid; e < kNB * kNB; e += 256) u[e & 63][e >> 6] = U[(e & 63) + (int64_t)(e >> 6) * ldu];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * 256 + tid;
    if (i >= m) return;
    float x[kNB];
#pragma unroll
    for (int j = 0; j < kNB; ++j) x[j] = X[i + j * ldx];
#pragma unroll
    for (int j = 0; j < kNB; ++j) {
        float v = x[j];
#pragma unroll
        for (int l = 0; l < j; ++l) v -= x[l] * u[l][j];
        x[j] = v * (1.0f / u[j][j]);
    }
#pragma unroll
    for (int j = 0; j < kNB; ++j) X[i + j * ldx] = x[j];
}

// Columns: Y(:, j) <- L^-1 Y(:, j) for a 64 x 64 unit lower L (LLNU).
__global__ void __launch_bounds__(256)
k_trsm_cols_llnu64(int64_t n, const float *__restrict__ L, int64_t ldl, float *__restrict__ Y,
                   int64_t ldy) {
    __shared__ float l[kNB][kNB + 1];
    const int tid = threadIdx.x;
    for (int e = tid; e < kNB * kNB; e += 256) l[e & 63][e >> 6] = L[(e & 63) + (int64_t)(e >> 6) * ldl];
    __syncthreads();
    const int64_t j = (int64_t)blockIdx.x * 256 + tid;
    if (j >= n) return;
    float y[kNB];
    const float *col = Y + j * ldy;
#pragma unroll
    for (int i = 0; i < kNB; ++i) y[i] = col[i];
#pragma unroll
    for (int c = 0; c < kNB; ++c) {
        const float t = y[c];
#pragma unroll
        for (int i = c + 1; i < kNB; ++i) y[i] -= t * l[i][c];
    }
    float *out = Y + j * ldy;
#pragma unroll
    for (int i = 0; i < kNB; ++i) out[i] = y[i];
}

// Rows: X(i, :) <- X(i, :) * L^-1 for a 64 x 64 unit lower L (RLNU),
// columns from the last to the first.
__global__ void __launch_bounds__(256)
k_trsm_rows_rlnu64(int64_t m, const float *__restrict__ L, int64_t ldl, float *__restrict__ X,
                   int64_t ldx) {
    __shared__ float l[kNB][kNB + 1];
    const int tid = threadIdx.x;
    for (int e = tid; e < kNB * kNB; e += 256) l[e & 63][e >> 6] = L[(e & 63) + (int64_t)(e >> 6) * ldl];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * 256 + tid;
    if (i >= m) return;
    float x[kNB];
#pragma unroll
    for (int j = 0; j < kNB; ++j) x[j] = X[i + j * ldx];
#pragma unroll
    for (int j = kNB - 1; j >= 0; --j) {
        float v = x[j];
#pragma unroll
        for (int q = j + 1; q < kNB; ++q) v -= l[q][j] * x[q];
        x[j] = v;
    }
#pragma unroll
    for (int j = 0; j < kNB; ++j) X[i + j * ldx] = x[j];
}

__global__ void k_identity(int64_t n, float *__restrict__ A, int64_t lda) {
    const int64_t j = blockIdx.y;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        A[i + j * lda] = (i == j) ? 1.0f : 0.0f;
}

__global__ void k_copy_block(int64_t m, const float *__restrict__ A, int64_t lda, float *__restrict__ B,
                             int64_t ldb) {
    const int64_t j = blockIdx.y;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
        B[i + j * ldb] = A[i + j * lda];
}

__global__ void k_set_int(int *p, int v) { *p = v; }

int sgemm(int64_t m, int64_t n, int64_t k, float alpha, const float *A, int64_t lda, const float *B,
          int64_t ldb, float beta, float *C, int64_t ldc, hipStream_t st) {
    if (m <= 0 || n <= 0) return OMX_OK;
    dim3 grid((unsigned)omx_cdiv(m, 64), (unsigned)omx_cdiv(n, 64));
    if (grid.y > 65535) return OMX_ERR_ARG;
    hipLaunchKernelGGL(k_sgemm_nn, grid, dim3(256), 0, st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

// sgemm with n split into chunks of at most 65535*64 columns
int sgemm_big(int64_t m, int64_t n, int64_t k, float alpha, const float *A, int64_t lda, const float *B,
              int64_t ldb, float beta, float *C, int64_t ldc, hipStream_t st) {
    const int64_t nmax = (int64_t)65535 * 64;
    for (int64_t j = 0; j < n; j += nmax) {
        const int64_t nn = (n - j < nmax) ? n - j : nmax;
        int rc = sgemm(m, nn, k, alpha, A, lda, B + j * ldb, ldb, beta, C + j * ldc, ldc, st);
        if (rc) return rc;
    }
    return OMX_OK;
}

}  // namespace

extern "C" {

int omx_sgemm_f32(int64_t m, int64_t n, int64_t k, float alpha, const float *A, int64_t lda,
                  const float *B, int64_t ldb, float beta, float *C, int64_t ldc, void *stream) {
    if (m < 0 || n < 0 || k <= 0 || k % 16 || !A || !B || !C || lda < m || ldb < k || ldc < m)
        return OMX_ERR_ARG;
    return sgemm_big(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, omx_stream(stream));
}

int omx_getrf_nopiv_f32(int64_t n, float *A, int64_t lda, int *info, void *stream) {
    if (n <= 0 || n % kNB || !A || lda < n) return OMX_ERR_ARG;
    hipStream_t st = omx_stream(stream);
    if (info) {
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, info, 0x7fffffff);
        OMX_LAUNCH_CHECK();
    }
    for (int64_t j = 0; j < n; j += kNB) {
        float *Ajj = A + j + j * lda;
        hipLaunchKernelGGL(k_getrf_diag64, dim3(1), dim3(256), 0, st, Ajj, lda, j, info);
        OMX_LAUNCH_CHECK();
        const int64_t rest = n - j - kNB;
        if (rest <= 0) break;
        hipLaunchKernelGGL(k_trsm_rows_runn64, dim3((unsigned)omx_cdiv(rest, 256)), dim3(256), 0, st, rest,
                           Ajj, lda, Ajj + kNB, lda);
        OMX_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_trsm_cols_llnu64, dim3((unsigned)omx_cdiv(rest, 256)), dim3(256), 0, st, rest,
                           Ajj, lda, Ajj + kNB * lda, lda);
        OMX_LAUNCH_CHECK();
        int rc = sgemm(rest, rest, kNB, -1.0f, Ajj + kNB, lda, Ajj + kNB * lda, lda, 1.0f,
                       Ajj + kNB + kNB * lda, lda, st);
        if (rc) return rc;
    }
    return OMX_OK;
}

int omx_trsm_f32(int variant, int64_t m, int64_t n, const float *T, int64_t ldt, float *B, int64_t ldb,
                 void *stream) {
    if (m < 0 || n < 0 || !T || !B || ldb < m) return OMX_ERR_ARG;
    if (m == 0 || n == 0) return OMX_OK;
    hipStream_t st = omx_stream(stream);
    if (variant == OMX_TRSM_RUNN || variant == OMX_TRSM_RLNU) {
        // B (m x n) times the inverse of an n x n triangle, n % 64 == 0
        if (n % kNB || ldt < n) return OMX_ERR_ARG;
        const unsigned g = (unsigned)omx_cdiv(m, 256);
        if (variant == OMX_TRSM_RUNN) {
            for (int64_t j = 0; j < n; j += kNB) {
                hipLaunchKernelGGL(k_trsm_rows_runn64, dim3(g), dim3(256), 0, st, m, T + j + j * ldt, ldt,
                                   B + j * ldb, ldb);
                OMX_LAUNCH_CHECK();
                const int64_t rest = n - j - kNB;
                if (rest > 0) {
                    int rc = sgemm_big(m, rest, kNB, -1.0f, B + j * ldb, ldb, T + j + (j + kNB) * ldt, ldt,
                                       1.0f, B + (j + kNB) * ldb, ldb, st);
                    if (rc) return rc;
                }
            }
        } else {
            for (int64_t j = n - kNB; j >= 0; j -= kNB) {
                hipLaunchKernelGGL(k_trsm_rows_rlnu64, dim3(g), dim3(256), 0, st, m, T + j + j * ldt, ldt,
                                   B + j * ldb, ldb);
                OMX_LAUNCH_CHECK();
                if (j > 0) {
                    int rc = sgemm_big(m, j, kNB, -1.0f, B + j * ldb, ldb, T + j, ldt, 1.0f, B, ldb, st);
                    if (rc) return rc;
                }
            }
        }
        return OMX_OK;
    }
    if (variant == OMX_TRSM_LLNU) {
        // inverse of an m x m unit lower triangle times B (m x n), m % 64 == 0
        if (m % kNB || ldt < m) return OMX_ERR_ARG;
        const int64_t nmax = (int64_t)65535 * 256;
        for (int64_t i = 0; i < m; i += kNB) {
            for (int64_t c = 0; c < n; c += nmax) {
                const int64_t nn = (n - c < nmax) ? n - c : nmax;
                hipLaunchKernelGGL(k_trsm_cols_llnu64, dim3((unsigned)omx_cdiv(nn, 256)), dim3(256), 0, st, nn,
                                   T + i + i * ldt, ldt, B + i + c * ldb, ldb);
                OMX_LAUNCH_CHECK();
            }
            const int64_t rest = m - i - kNB;
            if (rest > 0) {
                int rc = sgemm_big(rest, n, kNB, -1.0f, T + (i + kNB) + i * ldt, ldt, B + i, ldb, 1.0f,
                                   B + i + kNB, ldb, st);
                if (rc) return rc;
            }
        }
        return OMX_OK;
    }
    return OMX_ERR_ARG;
}

int omx_identity_f32(int64_t n, float *A, int64_t lda, void *stream) {
    if (n <= 0 || !A || lda < n || n > 65535) return OMX_ERR_ARG;
    hipLaunchKernelGGL(k_identity, dim3((unsigned)omx_cdiv(n, 256), (unsigned)n), dim3(256), 0,
                       omx_stream(stream), n, A, lda);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

int omx_copy_block_f32(int64_t m, int64_t n, const float *A, int64_t lda, float *B, int64_t ldb,
                       void *stream) {
    if (m < 0 || n < 0 || n > 65535 || !A || !B || lda < m || ldb < m) return OMX_ERR_ARG;
    if (m == 0 || n == 0) return OMX_OK;
    hipLaunchKernelGGL(k_copy_block, dim3((unsigned)omx_cdiv(m, 256), (unsigned)n), dim3(256), 0,
                       omx_stream(stream), m, A, lda, B, ldb);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

}  // extern "C"
