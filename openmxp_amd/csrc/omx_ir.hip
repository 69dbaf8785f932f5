// omx_ir.hip -- FP64 kernels of the iterative refinement: the on-the-fly
// residual GEMV (A regenerated from the LCG), infinity norms, vector ops and
// the blocked triangular solves with the FP32 LU factors.
#include <float.h>

#include <mutex>

#include "omx_common.hpp"

// Partial sums of omx_otf_gemv_f64_noalloc: the device memory that entry
// point uses in place of a caller's workspace.  64 MiB covers every order n
// (the residual GEMV's largest need is two column segments at n = 4194048,
// 67104768 bytes).  Each GPU the library runs on holds its own copy from
// module load until the process ends.  Host code never passes this name to a
// kernel (on the host it is only a stand-in for the symbol): it passes the
// device address hipGetSymbolAddress returns for the current device.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

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

// acc += A(i, j) x_j for the columns j of [j0, j1), in column order, the
// row's LCG state s stepping by the map jn per column (s leaves as the state
// of column j1).  Four columns per trip with a 32-bit trip count: the loop
// test and the state copies are then scalar or amortised, leaving per
// element the 64-bit affine step (four integer instructions), the exact
// int64 -> FP64 conversion (three), the scaling and the FMA.
__device__ __forceinline__ void otf_columns(uint64_t &s, double &acc, omx_map jn, const double *__restrict__ x,
                                            int64_t j0, int64_t j1) {
    const double *__restrict__ xp = x + j0;
    const int cnt = (int)(j1 - j0);
    int t = 0;
    for (; t + 4 <= cnt; t += 4) {
        const uint64_t s1 = omx_apply(jn, s);
        const uint64_t s2 = omx_apply(jn, s1);
        const uint64_t s3 = omx_apply(jn, s2);
        acc = __builtin_fma(omx_lcg_value(s), xp[t], acc);
        acc = __builtin_fma(omx_lcg_value(s1), xp[t + 1], acc);
        acc = __builtin_fma(omx_lcg_value(s2), xp[t + 2], acc);
        acc = __builtin_fma(omx_lcg_value(s3), xp[t + 3], acc);
        s = omx_apply(jn, s3);
    }
    for (; t < cnt; ++t) {
        acc = __builtin_fma(omx_lcg_value(s), xp[t], acc);
        s = omx_apply(jn, s);
    }
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
    // c0 <= lo <= hi <= c1
    const int64_t lo = c0 > d0 ? c0 : (d0 < c1 ? d0 : c1);
    const int64_t hi = c1 < d1 ? (c1 > lo ? c1 : lo) : (d1 > lo ? d1 : lo);
    otf_columns(s, acc, jn, x, c0, lo);
    const double di = diag[i];
    for (int64_t j = lo; j < hi; ++j) {
        const double a = (j == i) ? di : omx_lcg_value(s);
        acc = __builtin_fma(a, x[j], acc);
        s = omx_apply(jn, s);
    }
    otf_columns(s, acc, jn, x, hi, c1);
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
// block column J (solved values x_J final in x), workgroups of 1024
// threads (16 waves):
//   workgroup 0   rows of the next diagonal block D: r_D -= M(D, J) x_J, then
//                 D is solved in LDS: four sub-blocks of 64, each a 64-step
//                 lane-to-lane chain in one wave, while twelve more waves
//                 fold each solved sub-block into the later sub-blocks' rows;
//   workgroup w   the remaining rows beyond D, 256 per workgroup:
//                 r -= M(rows, J) x_J.
// A wave takes 16 of the 256 columns: its sixteen 16-byte loads per lane
// are in flight together, so a workgroup's 256 KB of the factor arrive in
// one round trip and the launch is bound by HBM rather than by the latency
// of one workgroup's reads.  Summation order per row and block column: each
// wave sums its 16 columns in order, the 16 partial sums are added in wave
// order, the total is subtracted from r.  Every FP32 factor entry is read
// once (2 n^2 bytes per triangle).
// ---------------------------------------------------------------------------
constexpr int kTB = 256;
constexpr int kSB = 64;
constexpr int kTW = 16;         // waves per workgroup
constexpr int kTT = 64 * kTW;   // threads per workgroup
constexpr int kTC = kTB / kTW;  // columns per wave

typedef float f32x4 __attribute__((ext_vector_type(4)));

struct TrsvShared {
    double xv[kTB];          // x_J
    double xs[kTB];          // r_D, then x_D
    double part[kTW][kTB];   // per-wave partial sums of rows_dot
    double pp[4][4][kSB];    // diagonal block: [sub-block][column chunk][row] sums
};

// Lane l's value of v in every lane (l uniform): two readlanes.
__device__ __forceinline__ double bcast_lane(double v, int l) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// sum over the block's ncols <= 256 columns from c0 of M(row, c) x_c for
// the 256 rows from r0 (rows >= rmax ignored); sh.xv holds x of the
// columns.  Wave w takes columns [16 w, 16 w + 16), lane l rows 4l .. 4l+3
// (one 16-byte load per column when kVec).  Returns the row's total in
// thread t = row - r0 for t < 256 (0 elsewhere).
template <bool kVec>
__device__ __forceinline__ double rows_dot(const float *__restrict__ A, int64_t lda, int64_t r0, int64_t rmax,
                                           int64_t c0, int ncols, TrsvShared &sh) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int cb = wave * kTC;
    const int ce = min(ncols, cb + kTC);
    const int64_t row = r0 + 4 * lane;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    if (kVec && row + 3 < rmax) {
        const float *p = A + row + (c0 + cb) * lda;
        if (ce - cb == kTC) {
            f32x4 v[kTC];
#pragma unroll
            for (int u = 0; u < kTC; ++u) v[u] = *reinterpret_cast<const f32x4 *>(p + (int64_t)u * lda);
#pragma unroll
            for (int u = 0; u < kTC; ++u) {
                const double xc = sh.xv[cb + u];
                s0 = __builtin_fma((double)v[u].x, xc, s0);
                s1 = __builtin_fma((double)v[u].y, xc, s1);
                s2 = __builtin_fma((double)v[u].z, xc, s2);
                s3 = __builtin_fma((double)v[u].w, xc, s3);
            }
        } else {
            for (int c = cb; c < ce; ++c) {
                const f32x4 v = *reinterpret_cast<const f32x4 *>(p + (int64_t)(c - cb) * lda);
                const double xc = sh.xv[c];
                s0 = __builtin_fma((double)v.x, xc, s0);
                s1 = __builtin_fma((double)v.y, xc, s1);
                s2 = __builtin_fma((double)v.z, xc, s2);
                s3 = __builtin_fma((double)v.w, xc, s3);
            }
        }
    } else {
        for (int c = cb; c < ce; ++c) {
            const float *q = A + row + (c0 + c) * lda;
            const double xc = sh.xv[c];
            if (row < rmax) s0 = __builtin_fma((double)q[0], xc, s0);
            if (row + 1 < rmax) s1 = __builtin_fma((double)q[1], xc, s1);
            if (row + 2 < rmax) s2 = __builtin_fma((double)q[2], xc, s2);
            if (row + 3 < rmax) s3 = __builtin_fma((double)q[3], xc, s3);
        }
    }
    sh.part[wave][4 * lane] = s0;
    sh.part[wave][4 * lane + 1] = s1;
    sh.part[wave][4 * lane + 2] = s2;
    sh.part[wave][4 * lane + 3] = s3;
    __syncthreads();
    double t = 0.0;
    if (tid < kTB) {
        t = sh.part[0][tid];
#pragma unroll
        for (int w = 1; w < kTW; ++w) t += sh.part[w][tid];
    }
    __syncthreads();
    return t;
}

// x_D <- T_DD^-1 r_D for the diagonal block at d0 of order dim (<= 256),
// unit lower (kUpper false, sub-blocks first to last) or upper non-unit
// (kUpper, last to first); r_D in sh.xs.  Roles:
//   waves 0-3    wave s runs sub-block s's 64-step chain at stage s (lane l:
//                row 64 s + l, its 64 entries of the diagonal sub-block in
//                registers, x_c read from lane c);
//   waves 4-15   wave 4 + 4 j + c accumulates, for the rows of target
//                sub-block t (lower t = j + 1, upper t = j), the products
//                with columns [16 c, 16 c + 16) of every sub-block solved
//                before t, each right after that sub-block's stage, and
//                hands its sum to t's chain, which subtracts the four chunk
//                sums (in chunk order) before it starts.
// All factor entries are loaded at the start, with row and column indices
// clamped into the block so every load is unconditional; entries outside
// the block meet x = 0 or are masked.  The upper chain divides by u_ii as
// q0 = r y, q = q0 + (r - u q0) y with y = RN(1/u): the correctly rounded
// quotient (Markstein's correction with a correctly rounded reciprocal).
template <bool kUpper>
__device__ __forceinline__ void diag_solve(const float *__restrict__ A, int64_t lda, int64_t d0, int dim,
                                           TrsvShared &sh) {
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nsb = (dim + kSB - 1) / kSB;
    const float *D = A + d0 + d0 * lda;
    const bool chain = wave < 4;
    const int t = chain ? wave : (kUpper ? (wave - 4) / 4 : 1 + (wave - 4) / 4);
    const int c = (wave - 4) & 3;
    const int i = kSB * t + lane;
    const float *Drow = D + min(i, dim - 1);
    float w[kSB];
    float ud = 1.0f;
    if (t < nsb) {
        if (chain) {
#pragma unroll
            for (int q = 0; q < kSB; ++q) w[q] = Drow[(int64_t)min(kSB * t + q, dim - 1) * lda];
            if (kUpper) ud = Drow[(int64_t)min(i, dim - 1) * lda];
            // lower: multipliers outside the triangle or past the block are
            // zeroed now, so that the chain's dependent steps carry no
            // selects (in the upper solve this spills; it selects in the
            // chain)
            if (!kUpper) {
#pragma unroll
                for (int q = 0; q < kSB; ++q) w[q] = (lane > q && i < dim) ? w[q] : 0.0f;
            }
        } else {
            // source sub-block s at slot s (lower, s < t) or s - 1 (upper, s > t)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int s = kUpper ? k + 1 : k;
                if (kUpper ? (s > t && s < nsb) : s < t) {
#pragma unroll
                    for (int u = 0; u < kTC; ++u)
                        w[kTC * k + u] = Drow[(int64_t)min(kSB * s + kTC * c + u, dim - 1) * lda];
                }
            }
        }
    }
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int st = kUpper ? 3 - k : k;
        if (st >= nsb) continue;
        if (chain && st == t) {
            const int c0 = kSB * st;
            const bool fed = kUpper ? st < nsb - 1 : st > 0;
            double v = (i < dim) ? sh.xs[i] : 0.0;
            if (fed) v -= ((sh.pp[st][0][lane] + sh.pp[st][1][lane]) + sh.pp[st][2][lane]) + sh.pp[st][3][lane];
            if (kUpper) {
                // the multipliers outside the triangle or past the block are
                // zeroed in place first (so the chain's steps carry no
                // selects; masking them with the loads spills here)
#pragma unroll
                for (int q = 0; q < kSB; ++q) {
                    float wq = (lane < q && c0 + q < dim) ? w[q] : 0.0f;
                    asm volatile("" : "+v"(wq));
                    w[q] = wq;
                }
                const double u = (i < dim) ? (double)ud : 1.0;
                const double y = 1.0 / u;
                auto quot = [&](double r) {
                    const double q0 = r * y;
                    return __builtin_fma(__builtin_fma(-u, q0, r), y, q0);
                };
                // step q: every lane forms its quotient, lane q's is x_q;
                // rows above subtract x_q u(i, q).  Lane q's own v is
                // untouched from then on, so its final quotient equals x_q.
#pragma unroll
                for (int q = kSB - 1; q >= 0; --q) v = __builtin_fma(-bcast_lane(quot(v), q), (double)w[q], v);
                v = quot(v);
            } else {
#pragma unroll
                for (int q = 0; q < kSB; ++q) {
                    // converted in place (hoisting all 64 conversions would
                    // take 128 registers)
                    float wq = w[q];
                    asm volatile("" : "+v"(wq));
                    v = __builtin_fma(-bcast_lane(v, q), (double)wq, v);
                }
            }
            if (i < dim) sh.xs[i] = v;
        }
        __syncthreads();
        if (!chain && t < nsb && (kUpper ? st > t : st < t)) {
            const int slot = kUpper ? st - 1 : st;
            const int cb = kSB * st + kTC * c;
            // four entries at a time: the scheduler must not convert all 16
            // up front (that would spill in the upper solve)
#pragma unroll
            for (int u = 0; u < kTC; ++u) {
                acc = __builtin_fma((double)w[kTC * slot + u], sh.xs[cb + u], acc);
                if ((u & 3) == 3) __builtin_amdgcn_sched_barrier(0);
            }
            if (st == (kUpper ? t + 1 : t - 1)) sh.pp[t][c][lane] = acc;
        }
        __syncthreads();
    }
}

// One block column of the forward solve.  j0 = first column of the solved
// block J (j0 < 0: nothing solved yet, only the first diagonal block).
template <bool kVec>
__global__ void __launch_bounds__(kTT)
k_trsv_lower_step(int64_t n, int64_t j0, const float *__restrict__ A, int64_t lda, double *__restrict__ x) {
    __shared__ TrsvShared sh;
    const int tid = threadIdx.x;
    const int64_t jb = j0 < 0 ? 0 : j0;
    const int ncols = j0 < 0 ? 0 : (int)min((int64_t)kTB, n - j0);
    const int64_t d0 = j0 < 0 ? 0 : j0 + kTB;
    if (ncols > 0 && tid < kTB) sh.xv[tid] = tid < ncols ? x[jb + tid] : 0.0;
    __syncthreads();
    if (blockIdx.x == 0) {
        const int dim = (int)min((int64_t)kTB, n - d0);
        const double t = ncols > 0 ? rows_dot<kVec>(A, lda, d0, d0 + dim, jb, ncols, sh) : 0.0;
        if (tid < kTB) sh.xs[tid] = (tid < dim) ? x[d0 + tid] - t : 0.0;
        __syncthreads();
        diag_solve<false>(A, lda, d0, dim, sh);
        if (tid < dim) x[d0 + tid] = sh.xs[tid];
        return;
    }
    const int64_t r0 = d0 + (int64_t)blockIdx.x * kTB;
    const double t = rows_dot<kVec>(A, lda, r0, n, jb, ncols, sh);
    if (tid < kTB && r0 + tid < n) x[r0 + tid] -= t;
}

// One block column of the backward solve: J = [j0, j0 + ncols) solved
// (j0 >= n: nothing solved yet, only the last diagonal block, which may be
// shorter than kTB); D = the block before J; rows [0, d0) updated.
template <bool kVec>
__global__ void __launch_bounds__(kTT)
k_trsv_upper_step(int64_t n, int64_t j0, const float *__restrict__ A, int64_t lda, double *__restrict__ x) {
    __shared__ TrsvShared sh;
    const int tid = threadIdx.x;
    const int64_t last0 = ((n - 1) / kTB) * kTB;  // first row of the last block
    const bool none = j0 >= n;
    const int ncols = none ? 0 : (int)min((int64_t)kTB, n - j0);
    const int64_t d0 = none ? last0 : j0 - kTB;
    const int dim = (int)min((int64_t)kTB, n - d0);
    if (ncols > 0 && tid < kTB) sh.xv[tid] = tid < ncols ? x[j0 + tid] : 0.0;
    __syncthreads();
    if (blockIdx.x == 0) {
        const double t = ncols > 0 ? rows_dot<kVec>(A, lda, d0, d0 + dim, j0, ncols, sh) : 0.0;
        if (tid < kTB) sh.xs[tid] = (tid < dim) ? x[d0 + tid] - t : 0.0;
        __syncthreads();
        diag_solve<true>(A, lda, d0, dim, sh);
        if (tid < dim) x[d0 + tid] = sh.xs[tid];
        return;
    }
    const int64_t r0 = (int64_t)(blockIdx.x - 1) * kTB;
    const double t = rows_dot<kVec>(A, lda, r0, d0, j0, ncols, sh);
    if (tid < kTB && r0 + tid < d0) x[r0 + tid] -= t;
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

}  // extern "C"

namespace {

// The residual GEMV's launches; part holds omx_otf_gemv_workspace_bytes(n).
int otf_gemv_launch(int64_t n, uint64_t seed, const double *diag, double alpha, const double *x, double *y,
                    void *work, hipStream_t st) {
    const int nseg = gemv_segments(n);
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

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (hipGetSymbolAddress with that device current) and kept.
int noalloc_scratch(double **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return OMX_ERR_HIP;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return OMX_ERR_HIP;
        g_scratch_dev[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_dev[dev];
    return OMX_OK;
}

}  // namespace

extern "C" {

int omx_otf_gemv_f64(int64_t n, uint64_t seed, const double *diag, double alpha,
                     const double *x, double *y, void *work, size_t work_bytes,
                     void *stream) {
    if (n <= 0 || !diag || !x || !y) return OMX_ERR_ARG;
    const size_t need = omx_otf_gemv_workspace_bytes(n);
    if (need > 0 && (!work || work_bytes < need)) return OMX_ERR_WORKSPACE;
    return otf_gemv_launch(n, seed, diag, alpha, x, y, work, omx_stream(stream));
}

int omx_otf_gemv_f64_noalloc(int64_t n, uint64_t seed, const double *diag, double alpha, const double *x,
                             double *y, void *stream) {
    if (n <= 0 || !diag || !x || !y) return OMX_ERR_ARG;
    if (omx_otf_gemv_workspace_bytes(n) > kNoAllocScratchBytes) return OMX_ERR_WORKSPACE;
    double *part = nullptr;
    const int rc = noalloc_scratch(&part);
    if (rc) return rc;
    return otf_gemv_launch(n, seed, diag, alpha, x, y, part, omx_stream(stream));
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
            hipLaunchKernelGGL(k_trsv_lower_step<true>, dim3(grid), dim3(kTT), 0, st, n, j0, A, lda, x);
        else
            hipLaunchKernelGGL(k_trsv_lower_step<false>, dim3(grid), dim3(kTT), 0, st, n, j0, A, lda, x);
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
            hipLaunchKernelGGL(k_trsv_upper_step<true>, dim3(grid), dim3(kTT), 0, st, n, j0, A, lda, x);
        else
            hipLaunchKernelGGL(k_trsv_upper_step<false>, dim3(grid), dim3(kTT), 0, st, n, j0, A, lda, x);
        OMX_LAUNCH_CHECK();
    }
    return OMX_OK;
}

double omx_harness(double normr, double norma, double normx, double normb, int64_t n) {
    return normr / (norma * normx + normb) * 1.0 / ((double)n * DBL_EPSILON / 2);
}

}  // extern "C"
