// omx_matgen.hip -- on-device LCG generators for the system matrix, its
// diagonal and the right-hand side (bit-exact with the reference's
// hpl_rand.hpp / matgen.hpp).
//
// Layout in HBM: A is column-major FP32 with leading dimension lda (the
// order-n system in one allocation, 4*n*lda bytes); diag and b are FP64
// vectors of length n.
#include "omx_common.hpp"

// The diagonal must be summed exactly as the reference does, so no
// floating-point contraction may change the rounding in this file.
#pragma clang fp contract(off)

namespace {

constexpr int kFillThreads = 256;
constexpr int kDiagThreads = 512;  // MATGEM_THREADS, matgen.hpp:7

// One workgroup per column (grid-stride over columns); thread t writes rows
// t, t+256, ... stepping the LCG by a^256 per row group, so every store of a
// wave is a contiguous 256-byte run.  fill_panel_dev (matgen.hpp:67-82)
// walks rows per block instead; the values are identical.
__global__ void __launch_bounds__(kFillThreads)
k_matgen_block(int64_t n, uint64_t x00, int64_t i0, int64_t j0, int64_t m,
               int64_t k, float *__restrict__ A, int64_t lda, omx_map step) {
    const int64_t t = threadIdx.x;
    if (t >= m) return;
    for (int64_t j = blockIdx.x; j < k; j += gridDim.x) {
        uint64_t s = omx_apply(omx_jump((uint64_t)(i0 + t) + (uint64_t)n * (uint64_t)(j0 + j)), x00);
        float *col = A + j * lda;
        for (int64_t i = t; i < m; i += kFillThreads) {
            col[i] = (float)omx_lcg_value(s);
            s = omx_apply(step, s);
        }
    }
}

// One workgroup of 512 threads per row.  Thread `id` sums |A_gj| for
// j = id, id+512, ... (j != g) in order, then the tree 512 -> 4 and
// ((s0+s1)+s2)+s3, exactly as fill_panel_diag_dev (matgen.hpp:26-61).
// Optionally stores A_gg = float(d) (matgen.hpp:62).
__global__ void __launch_bounds__(kDiagThreads)
k_diag(int64_t n, uint64_t x00, int64_t i0, double *__restrict__ diag,
       float *__restrict__ A, int64_t lda, omx_map jump_thread) {
    __shared__ double sh[kDiagThreads];
    const int id = threadIdx.x;
    const int64_t r = blockIdx.x;
    const int64_t g = i0 + r;
    uint64_t s = omx_apply(omx_jump((uint64_t)g + (uint64_t)n * (uint64_t)id), x00);
    double t = 0.0;
    for (int64_t j = id; j < n; j += kDiagThreads) {
        if (j != g) t += fabs(omx_lcg_value(s));
        s = omx_apply(jump_thread, s);
    }
    sh[id] = t;
    __syncthreads();
#pragma unroll
    for (int w = kDiagThreads / 2; w >= 4; w >>= 1) {
        if (id < w) sh[id] += sh[id + w];
        __syncthreads();
    }
    if (id == 0) {
        const double d = sh[0] + sh[1] + sh[2] + sh[3];
        diag[r] = d;
        if (A) A[g + g * lda] = (float)d;
    }
}

__global__ void k_rhs(int64_t n, uint64_t x_first, double *__restrict__ b, omx_map step) {
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    if (i0 >= n) return;
    uint64_t s = omx_apply(omx_jump((uint64_t)i0), x_first);
    for (int64_t i = i0; i < n; i += stride) {
        b[i] = omx_lcg_value(s);
        s = omx_apply(step, s);
    }
}

int launch_block(int64_t n, uint64_t seed, int64_t i0, int64_t j0, int64_t m,
                 int64_t k, float *A, int64_t lda, hipStream_t st) {
    if (m <= 0 || k <= 0) return OMX_OK;
    const int64_t grid = k < (1 << 20) ? k : (1 << 20);
    hipLaunchKernelGGL(k_matgen_block, dim3((unsigned)grid), dim3(kFillThreads), 0, st, n,
                       omx_lcg_init(seed), i0, j0, m, k, A, lda, omx_jump(kFillThreads));
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

int launch_diag(int64_t n, uint64_t seed, int64_t i0, int64_t m, double *diag,
                float *A, int64_t lda, hipStream_t st) {
    if (m <= 0) return OMX_OK;
    hipLaunchKernelGGL(k_diag, dim3((unsigned)m), dim3(kDiagThreads), 0, st, n,
                       omx_lcg_init(seed), i0, diag, A, lda,
                       omx_jump((uint64_t)kDiagThreads * (uint64_t)n));
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

}  // namespace

extern "C" {

const char *omx_version(void) { return "omxp 0.1 (gfx950)"; }

int omx_matgen_f32(int64_t n, uint64_t seed, float *A, int64_t lda, double *diag,
                   void *stream) {
    if (n <= 0 || lda < n || !A || !diag || n > ((int64_t)1 << 31)) return OMX_ERR_ARG;
    hipStream_t st = omx_stream(stream);
    int rc = launch_block(n, seed, 0, 0, n, n, A, lda, st);
    if (rc) return rc;
    return launch_diag(n, seed, 0, n, diag, A, lda, st);
}

int omx_matgen_block_f32(int64_t n, uint64_t seed, int64_t i0, int64_t j0, int64_t m,
                         int64_t k, float *A, int64_t lda, void *stream) {
    if (n <= 0 || i0 < 0 || j0 < 0 || m < 0 || k < 0 || lda < m || (!A && m * k > 0))
        return OMX_ERR_ARG;
    return launch_block(n, seed, i0, j0, m, k, A, lda, omx_stream(stream));
}

int omx_diag_f64(int64_t n, uint64_t seed, int64_t i0, int64_t m, double *diag,
                 void *stream) {
    if (n <= 0 || i0 < 0 || m < 0 || i0 + m > n || (!diag && m > 0)) return OMX_ERR_ARG;
    return launch_diag(n, seed, i0, m, diag, nullptr, 0, omx_stream(stream));
}

int omx_rhs_f64(int64_t n, uint64_t seed, double *b, void *stream) {
    if (n <= 0 || !b) return OMX_ERR_ARG;
    const int threads = 256;
    int64_t blocks = omx_cdiv(n, threads * 16);
    if (blocks < 1) blocks = 1;
    const uint64_t x_first =
        omx_apply(omx_jump((uint64_t)n * (uint64_t)n), omx_lcg_init(seed));
    hipLaunchKernelGGL(k_rhs, dim3((unsigned)blocks), dim3(threads), 0, omx_stream(stream), n,
                       x_first, b, omx_jump((uint64_t)(blocks * threads)));
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

}  // extern "C"
