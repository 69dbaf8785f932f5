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
#include <algorithm>
#include <utility>
#include <vector>

#include "omx_common.hpp"

namespace {

constexpr int kNB = 64;
#define RC(x)                  \
    do {                       \
        const int rc_ = (x);   \
        if (rc_) return rc_;   \
    } while (0)

// The triangular solves (omx_trsm_f32) go 64 columns at a time inside
// blocks of kOB: the 64-wide steps update only the rest of their block, and
// each block updates everything beyond it in one product of depth kOB, so
// the large products read and write their C a quarter as often.
constexpr int64_t kOB = 256;

// ---------------------------------------------------------------------------
// FP32 GEMM, C = beta*C + alpha * A * B (all column-major, no transposes).
// 64 x 64 tile per workgroup of 256 threads, 4 x 4 results per thread,
// k advanced 16 at a time through LDS.  k must be a multiple of 16.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_sgemm_nn(int64_t M, int64_t N, int64_t K, float alpha, const float *__restrict__ A, int64_t lda,
           const float *__restrict__ B, int64_t ldb, float beta, float *__restrict__ C, int64_t ldc,
           int64_t sA, int64_t sB, int64_t sC) {
    __shared__ float As[16][64 + 4];
    __shared__ float Bs[16][64 + 4];
    // batch blockIdx.z: operands sA / sB / sC floats further on
    A += (int64_t)blockIdx.z * sA;
    B += (int64_t)blockIdx.z * sB;
    C += (int64_t)blockIdx.z * sC;
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
// The same product on the matrix cores (v_mfma_f32_32x32x2_f32), for the
// large updates: a kT x kT tile (128, or 64 where 128-tiles would leave CUs
// idle) per workgroup of four waves (a kT/2 square quadrant each, made of
// 32 x 32 MFMA blocks), k staged through LDS up to 64 at a time, the next stage's global loads in flight during the current
// stage's MFMAs.  The MFMA's first operand takes B (lane l: column l % 32),
// the second A (lane l: row l % 32), so every accumulator register of lane l
// belongs to row l % 32 of its block and a half-wave's C access is 32
// consecutive rows of one column (one 128-byte line).  In a stage of kc
// columns, lanes 0-31 carry k in [0, kc/2) and lanes 32-63 k in [kc/2, kc);
// the MFMA adds the two halves.  Needs A and B 16-byte aligned with lda and
// ldb multiples of 4, and k % 16 == 0 (so kc is a multiple of 16).
// ---------------------------------------------------------------------------
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int kMK = 64;         // k per LDS stage
constexpr int kBLd = kMK + 4;   // row stride of Bs: conflict-free 16-byte reads

template <int kT, bool kOneStage>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_sgemm_mfma(int64_t M, int64_t N, int64_t K, float alpha, const float *__restrict__ A, int64_t lda,
             const float *__restrict__ B, int64_t ldb, float beta, float *__restrict__ C, int64_t ldc,
             int64_t sA, int64_t sB, int64_t sC) {
    constexpr int kB = kT / 64;       // MFMA blocks per quadrant side
    constexpr int kS = kT / 16;       // float4 loads per thread per operand and stage
    constexpr int kAK = 1024 / kT;    // stage columns between a thread's A loads
    __shared__ float As[kMK][kT];     // As[k][i] = A(i0 + i, k0 + k)
    __shared__ float Bs[kT][kBLd];    // Bs[j][k] = B(k0 + k, j0 + j)
    A += (int64_t)blockIdx.z * sA;
    B += (int64_t)blockIdx.z * sB;
    C += (int64_t)blockIdx.z * sC;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l32 = lane & 31, h = lane >> 5;
    const int64_t i0 = (int64_t)blockIdx.x * kT, j0 = (int64_t)blockIdx.y * kT;
    const int wi = (wave & 1) * (kT / 2), wj = (wave >> 1) * (kT / 2);
    // global -> LDS: A rows ar .. ar+3 at stage columns ak + kAK s; B column
    // bj + 16 s, stage rows bk .. bk+3.  Rows (columns) past M (N) are
    // clamped: their products land only in results that are not stored.
    const int ar = 4 * (tid % (kT / 4)), ak = tid / (kT / 4);
    const int bk = 4 * (tid & 15), bj = tid >> 4;
    const int64_t arow = i0 + ar;
    const bool afull = arow + 3 < M;
    int64_t ar1 = arow + 1, ar2 = arow + 2, ar3 = arow + 3;
    const int64_t arow0 = arow < M ? arow : M - 1;
    if (ar1 >= M) ar1 = M - 1;
    if (ar2 >= M) ar2 = M - 1;
    if (ar3 >= M) ar3 = M - 1;
    const float *Bcol[kS];
#pragma unroll
    for (int s = 0; s < kS; ++s) {
        int64_t j = j0 + bj + 16 * s;
        if (j >= N) j = N - 1;
        Bcol[s] = B + j * ldb;
    }
    f32x4 ra[kS], rb[kS];
    // stage k0 .. k0 + kc into ra / rb.  Every load is unconditional: stage
    // columns (rows) at or past kc read column (row) kc - 1 (kc - 4) instead,
    // and are not stored to LDS.
#define OMX_SGEMM_GLOAD(k0, kc)                                                                            \
    {                                                                                                      \
        _Pragma("unroll") for (int s = 0; s < kS; ++s) {                                                   \
            const int kk = ak + kAK * s < (kc) ? ak + kAK * s : (kc) - 1;                                  \
            const float *p = A + ((k0) + kk) * lda;                                                        \
            ra[s] = afull ? *reinterpret_cast<const f32x4 *>(p + arow)                                    \
                          : f32x4{p[arow0], p[ar1], p[ar2], p[ar3]};                                 \
        }                                                                                                  \
        const int kb4 = bk < (kc) ? bk : (kc) - 4;                                                         \
        _Pragma("unroll") for (int s = 0; s < kS; ++s)                                                     \
            rb[s] = *reinterpret_cast<const f32x4 *>(Bcol[s] + (k0) + kb4);                               \
    }
    f32x16 acc[kB][kB];
#pragma unroll
    for (int a = 0; a < kB; ++a)
#pragma unroll
        for (int b = 0; b < kB; ++b)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.0f;
    int64_t k0 = 0;
    int kc = (int)(K < kMK ? K : kMK);
    OMX_SGEMM_GLOAD(0, kc);
    // kOneStage (K <= 64): C of the wave's quadrant is read now, with the
    // operands, so that its latency overlaps theirs and the MFMAs (rows /
    // columns past the edge read a clamped entry and are not stored);
    // cv[a][b][v] sits where acc[a][b][v] does (see below).  With several
    // stages the staging registers are live through the loop and C is read
    // in the epilogue.
    float cv[kB][kB][16];
    if (kOneStage && beta != 0.0f) {
#pragma unroll
        for (int a = 0; a < kB; ++a) {
            int64_t i = i0 + wi + 32 * a + l32;
            if (i >= M) i = M - 1;
#pragma unroll
            for (int b = 0; b < kB; ++b)
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    int64_t j = j0 + wj + 32 * b + 4 * h + 8 * (v >> 2) + (v & 3);
                    if (j >= N) j = N - 1;
                    cv[a][b][v] = C[i + j * ldc];
                }
        }
    }
    for (;;) {
#pragma unroll
        for (int s = 0; s < kS; ++s)
            if (ak + kAK * s < kc) *reinterpret_cast<f32x4 *>(&As[ak + kAK * s][ar]) = ra[s];
        if (bk < kc) {
#pragma unroll
            for (int s = 0; s < kS; ++s) *reinterpret_cast<f32x4 *>(&Bs[bj + 16 * s][bk]) = rb[s];
        }
        __syncthreads();
        const int64_t kn = k0 + kc;
        const int kcn = (!kOneStage && kn < K) ? (int)(K - kn < kMK ? K - kn : kMK) : 0;
        if (kcn) OMX_SGEMM_GLOAD(kn, kcn);
        const int half = kc >> 1, kb = h * half;
        for (int s = 0; s < half; s += 4) {
            f32x4 y[kB];
#pragma unroll
            for (int b = 0; b < kB; ++b) y[b] = *reinterpret_cast<const f32x4 *>(&Bs[wj + 32 * b + l32][kb + s]);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                float x[kB];
#pragma unroll
                for (int a = 0; a < kB; ++a) x[a] = As[kb + s + u][wi + 32 * a + l32];
#pragma unroll
                for (int a = 0; a < kB; ++a)
#pragma unroll
                    for (int b = 0; b < kB; ++b)
                        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(y[b][u], x[a], acc[a][b], 0, 0, 0);
            }
        }
        __syncthreads();
        if (!kcn) break;
        k0 = kn;
        kc = kcn;
    }
#undef OMX_SGEMM_GLOAD
    // acc[a][b][v]: row i0 + wi + 32 a + l32, column j0 + wj + 32 b + 8 (v / 4) + 4 h + v % 4
#pragma unroll
    for (int a = 0; a < kB; ++a) {
        const int64_t i = i0 + wi + 32 * a + l32;
        if (i >= M) continue;
#pragma unroll
        for (int b = 0; b < kB; ++b) {
            const int64_t jb = j0 + wj + 32 * b + 4 * h;
            if (!kOneStage && beta != 0.0f) {
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int64_t j = jb + 8 * (v >> 2) + (v & 3);
                    cv[a][b][v] = j < N ? C[i + j * ldc] : 0.0f;
                }
            }
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int64_t j = jb + 8 * (v >> 2) + (v & 3);
                if (j < N)
                    C[i + j * ldc] = (beta == 0.0f) ? alpha * acc[a][b][v] : beta * cv[a][b][v] + alpha * acc[a][b][v];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// 64 x 64 diagonal piece of the LU: unblocked right-looking LU in one wave.
// Lane i holds row i in registers.  Step k: every lane reads row k from
// lane k (v_readlane into scalar registers, no LDS round trip); the lanes
// below scale their column-k entry by 1/u_kk (the multiplier) and subtract
// multiplier x row k from the rest of their row.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float lane_bcast(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

// Lane i of the calling wave: row i of the 64 x 64 block in r, factored in
// place (multipliers below the diagonal, U on and above it).  info (may be
// null): the smallest row0 + k + 1 with a zero pivot u_kk.
__device__ __forceinline__ void lu64_wave(float (&r)[kNB], int i, int64_t row0, int *info) {
#pragma unroll
    for (int k = 0; k < kNB; ++k) {
        const float ukk = lane_bcast(r[k], k);
        if (i == 0 && ukk == 0.0f && info) atomicMin(info, (int)(row0 + k + 1));
        const float inv = 1.0f / ukk;
        const bool below = i > k;
        // rows at or above k take multiplier 0: fma(-0, u, r) leaves r as
        // it is (up to the sign of a zero), with no select per entry
        const float l = below ? r[k] * inv : 0.0f;
        if (below) r[k] = l;
#pragma unroll
        for (int j = k + 1; j < kNB; ++j) r[j] = __builtin_fmaf(-l, lane_bcast(r[j], k), r[j]);
    }
}

// The 64 x 64 triangle T (column-major, leading dimension ldt) into LDS:
// t[r][c] = T(r, c).
__device__ __forceinline__ void load_tri64(const float *__restrict__ T, int64_t ldt, float (*t)[kNB + 1]) {
    for (int e = threadIdx.x; e < kNB * kNB; e += blockDim.x) t[e & 63][e >> 6] = T[(e & 63) + (int64_t)(e >> 6) * ldt];
    __syncthreads();
}

// Solves of one row (RUNN) or one column (LLNU) of 64 entries by a quad of
// lanes: lane q of the quad owns entries q, q + 4, ..., q + 60 (16 registers,
// index s <-> entry q + 4 s).  Step j: the owner (lane j % 4) finalises
// entry j, the quad reads it with a DPP broadcast, and every lane updates
// its later entries.  Each entry still receives its terms in the order
// j = 0, 1, ... with one fused multiply-add each, as in column-by-column
// substitution, so the results are those of the one-thread sweep; the
// dependent chain per step is one multiply (or FMA) and one broadcast.

// rd[j] = 1 / t[j][j] (the diagonal's reciprocals, shared by the workgroup)
__device__ __forceinline__ void diag_recip64(const float (*t)[kNB + 1], float *rd) {
    if (threadIdx.x < kNB) rd[threadIdx.x] = 1.0f / t[threadIdx.x][threadIdx.x];
    __syncthreads();
}

// value of v in quad lane o (o compile-time) for every lane of the quad
template <int o>
__device__ __forceinline__ float quad_bcast(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), o | (o << 2) | (o << 4) | (o << 6),
                                                              0xF, 0xF, false));
}

template <int j>
__device__ __forceinline__ void runn_step(float (&x)[16], int q, const float (*u)[kNB + 1], const float *rd) {
    constexpr int o = j & 3, s = j >> 2;
    const float cand = x[s] * rd[j];
    const float xj = quad_bcast<o>(cand);
    if (q == o) x[s] = xj;
    if (q > o) x[s] = __builtin_fmaf(-xj, u[j][q + 4 * s], x[s]);
#pragma unroll
    for (int t = s + 1; t < 16; ++t) x[t] = __builtin_fmaf(-xj, u[j][q + 4 * t], x[t]);
}

template <int j>
__device__ __forceinline__ void llnu_step(float (&y)[16], int q, const float (*l)[kNB + 1]) {
    constexpr int o = j & 3, s = j >> 2;
    const float t = quad_bcast<o>(y[s]);
    if (q > o) y[s] = __builtin_fmaf(-t, l[q + 4 * s][j], y[s]);
#pragma unroll
    for (int w = s + 1; w < 16; ++w) y[w] = __builtin_fmaf(-t, l[q + 4 * w][j], y[w]);
}

template <int... J>
__device__ __forceinline__ void runn_all(float (&x)[16], int q, const float (*u)[kNB + 1], const float *rd,
                                         std::integer_sequence<int, J...>) {
    (runn_step<J>(x, q, u, rd), ...);
}

template <int... J>
__device__ __forceinline__ void llnu_all(float (&y)[16], int q, const float (*l)[kNB + 1],
                                         std::integer_sequence<int, J...>) {
    (llnu_step<J>(y, q, l), ...);
}

// Row i: X(i, :) <- X(i, :) * U^-1 for a 64 x 64 upper non-unit U (RUNN);
// x_j = (b_j - sum_{l<j} x_l u_lj) * rd_j with rd_j = 1 / u_jj (correctly
// rounded, formed once per workgroup).  Called by the four lanes of a quad
// with the same i (q = lane % 4).
__device__ __forceinline__ void row_runn64(int64_t i, int q, const float (*u)[kNB + 1], const float *rd,
                                           float *__restrict__ X, int64_t ldx) {
    float x[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) x[t] = X[i + (q + 4 * t) * ldx];
    runn_all(x, q, u, rd, std::make_integer_sequence<int, kNB>{});
#pragma unroll
    for (int t = 0; t < 16; ++t) X[i + (q + 4 * t) * ldx] = x[t];
}

// Column j: Y(:, j) <- L^-1 Y(:, j) for a 64 x 64 unit lower L (LLNU);
// y_i -= y_c l_ic for c < i, c ascending.  Called by a quad as above.
__device__ __forceinline__ void col_llnu64(int64_t j, int q, const float (*l)[kNB + 1], float *__restrict__ Y,
                                           int64_t ldy) {
    float y[16];
    float *col = Y + j * ldy;
#pragma unroll
    for (int t = 0; t < 16; ++t) y[t] = col[q + 4 * t];
    llnu_all(y, q, l, std::make_integer_sequence<int, kNB>{});
#pragma unroll
    for (int t = 0; t < 16; ++t) col[q + 4 * t] = y[t];
}

// 64 rows (or columns) per workgroup of 256 threads, a quad per row.  The
// quad's lanes stay together: a row past the end exits as a whole quad.
__global__ void __launch_bounds__(256)
k_trsm_rows_runn64(int64_t m, const float *__restrict__ U, int64_t ldu, float *__restrict__ X,
                   int64_t ldx) {
    __shared__ float u[kNB][kNB + 1];
    __shared__ float rd[kNB];
    load_tri64(U, ldu, u);
    diag_recip64(u, rd);
    const int64_t i = (int64_t)blockIdx.x * 64 + (threadIdx.x >> 2);
    if (i < m) row_runn64(i, threadIdx.x & 3, u, rd, X, ldx);
}

__global__ void __launch_bounds__(256)
k_trsm_cols_llnu64(int64_t n, const float *__restrict__ L, int64_t ldl, float *__restrict__ Y,
                   int64_t ldy) {
    __shared__ float l[kNB][kNB + 1];
    load_tri64(L, ldl, l);
    const int64_t j = (int64_t)blockIdx.x * 64 + (threadIdx.x >> 2);
    if (j < n) col_llnu64(j, threadIdx.x & 3, l, Y, ldy);
}

// One 64-column step of the blocked LU in one launch.  Every workgroup
// factors the 64 x 64 diagonal block D at Ajj itself (wave 0, lu64_wave;
// all get the same factors) into LDS; workgroup 0 reports zero pivots.
// Then workgroups [0, nrw) solve the rows below D (RUNN with D's U), the
// next nrw the columns right of D (LLNU with D's unit L), 64 per workgroup,
// a quad per row or column; rest = rows below = columns right of D.  Each
// quad's entries are loaded before the factorisation, so the loads overlap
// it.
// No workgroup writes D back in its own step: the workgroups of a launch
// need not run at the same time (a concurrent kernel may hold the CUs), so
// one that starts after D was overwritten would factor the factors.  The
// factored D of the previous step (still the original in memory, and read
// by no other workgroup of this launch) is written by one more workgroup,
// the last one (index max(1, 2 nrw)), when prev is set; the last step's
// workgroup 0, with nothing below or right of D, writes its own D.
__global__ void __launch_bounds__(256)
k_getrf_step64(int64_t rest, int nrw, float *__restrict__ Ajj, int64_t lda, int64_t row0, int *__restrict__ info,
               int prev) {
    __shared__ float d[kNB][kNB + 1];
    __shared__ float rd[kNB];
    const int tid = threadIdx.x, q = tid & 3;
    if (prev && (int)blockIdx.x == (nrw > 0 ? 2 * nrw : 1)) {
        // the previous step's diagonal block, 64 rows and columns up
        if (tid < kNB) {
            float *Dp = Ajj - kNB - kNB * lda;
            float r[kNB];
#pragma unroll
            for (int j = 0; j < kNB; ++j) r[j] = Dp[tid + (int64_t)j * lda];
            lu64_wave(r, tid, row0 - kNB, nullptr);
#pragma unroll
            for (int j = 0; j < kNB; ++j) Dp[tid + (int64_t)j * lda] = r[j];
        }
        return;
    }
    const bool rows = (int)blockIdx.x < nrw;
    const int64_t e = (int64_t)(rows ? (int)blockIdx.x : (int)blockIdx.x - nrw) * 64 + (tid >> 2);
    const bool act = e < rest;
    // row e below D: entries (e, q + 4 t); column e right of D: (q + 4 t, e)
    float *base = rows ? Ajj + kNB + e : Ajj + (kNB + e) * lda;
    const int64_t step = rows ? lda : 1;
    float x[16];
    if (act) {
#pragma unroll
        for (int t = 0; t < 16; ++t) x[t] = base[(q + 4 * t) * step];
    }
    if (tid < kNB) {
        float r[kNB];
#pragma unroll
        for (int j = 0; j < kNB; ++j) r[j] = Ajj[tid + (int64_t)j * lda];
        lu64_wave(r, tid, row0, blockIdx.x == 0 ? info : nullptr);
#pragma unroll
        for (int j = 0; j < kNB; ++j) d[tid][j] = r[j];
        if (nrw == 0) {
#pragma unroll
            for (int j = 0; j < kNB; ++j) Ajj[tid + (int64_t)j * lda] = r[j];
        }
    }
    __syncthreads();
    diag_recip64(d, rd);
    if (!act) return;
    if (rows)
        runn_all(x, q, d, rd, std::make_integer_sequence<int, kNB>{});
    else
        llnu_all(x, q, d, std::make_integer_sequence<int, kNB>{});
#pragma unroll
    for (int t = 0; t < 16; ++t) base[(q + 4 * t) * step] = x[t];
}

// Rows: X(i, :) <- X(i, :) * L^-1 for a 64 x 64 unit lower L (RLNU),
// columns from the last to the first.
__global__ void __launch_bounds__(256)
k_trsm_rows_rlnu64(int64_t m, const float *__restrict__ L, int64_t ldl, float *__restrict__ X,
                   int64_t ldx) {
    __shared__ float l[kNB][kNB + 1];
    load_tri64(L, ldl, l);
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
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

// Inverses of the 64 x 64 diagonal blocks of the factored block D (L unit
// lower, U upper in one array): workgroup w < nblk writes U_ww^-1 into Xu
// (row r of it = e_r U_ww^-1, a RUNN row solve of the identity row by a lane
// quad), workgroup nblk + w writes L_ww^-1 into Xl (column c = L_ww^-1 e_c,
// an LLNU column solve by a lane quad).
__global__ void __launch_bounds__(256)
k_tri_inv64(int nblk, const float *__restrict__ D, int64_t ldd, float *__restrict__ Xu, float *__restrict__ Xl,
            int64_t ldx) {
    __shared__ float t[kNB][kNB + 1];
    __shared__ float rd[kNB];
    const bool up = (int)blockIdx.x < nblk;
    const int64_t o = (int64_t)(up ? blockIdx.x : blockIdx.x - nblk) * kNB;
    load_tri64(D + o + o * ldd, ldd, t);
    const int q = threadIdx.x & 3, r = threadIdx.x >> 2;
    float v[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) v[s] = (q + 4 * s == r) ? 1.0f : 0.0f;
    if (up) {
        diag_recip64(t, rd);
        runn_all(v, q, t, rd, std::make_integer_sequence<int, kNB>{});
        float *row = Xu + o + r + o * ldx;
#pragma unroll
        for (int s = 0; s < 16; ++s) row[(q + 4 * s) * ldx] = v[s];
    } else {
        llnu_all(v, q, t, std::make_integer_sequence<int, kNB>{});
        float *col = Xl + o + (o + r) * ldx;
#pragma unroll
        for (int s = 0; s < 16; ++s) col[q + 4 * s] = v[s];
    }
}

// The entries of the two inverses that k_tri_inv64 and the merges do not
// write: in column j, Xu below and Xl above j's diagonal 64 x 64 block (the
// diagonal blocks are written whole, the merges write every off-diagonal
// block of the other side).  One workgroup per column.
__global__ void k_zero_offdiag(int64_t n, float *__restrict__ Xu, float *__restrict__ Xl, int64_t ld) {
    const int64_t j = blockIdx.x;
    const int64_t top = (j / kNB) * kNB, bot = top + kNB;
    float *u = Xu + j * ld, *l = Xl + j * ld;
    for (int64_t i = bot + threadIdx.x; i < n; i += blockDim.x) u[i] = 0.0f;
    for (int64_t i = threadIdx.x; i < top; i += blockDim.x) l[i] = 0.0f;
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

// The MFMA kernel takes products with at least this many C entries per
// launch (2^18 = 512 x 512; smaller ones are latency-bound and the 64 x 64
// VALU tiles spread them over more CUs) when its alignment conditions hold.
#ifndef OMX_SGEMM_MFMA_T64
#define OMX_SGEMM_MFMA_T64 1024
#endif
#ifndef OMX_SGEMM_MFMA_MIN
#define OMX_SGEMM_MFMA_MIN (int64_t{1} << 18)
#endif

bool use_mfma(int64_t m, int64_t n, int64_t k, int64_t batch, const float *A, int64_t lda, const float *B,
              int64_t ldb) {
    const bool aligned = (reinterpret_cast<uintptr_t>(A) & 15) == 0 && (reinterpret_cast<uintptr_t>(B) & 15) == 0 &&
                         (lda & 3) == 0 && (ldb & 3) == 0 && k >= 16 && k % 16 == 0;
    return aligned && m * n * batch >= OMX_SGEMM_MFMA_MIN;
}

int sgemm_mfma(int64_t m, int64_t n, int64_t k, float alpha, const float *A, int64_t lda, int64_t sA, const float *B,
               int64_t ldb, int64_t sB, float beta, float *C, int64_t ldc, int64_t sC, int64_t batch, hipStream_t st) {
    // 64-tiles where 128-tiles would give fewer than 1024 workgroups (four
    // per CU): the inverse merges (few tiles, deep k) and the mid-size LU
    // updates finish sooner on four times as many, smaller workgroups
    const bool small = omx_cdiv(m, 128) * omx_cdiv(n, 128) * batch < OMX_SGEMM_MFMA_T64;
    const int t = small ? 64 : 128;
    const dim3 grid((unsigned)omx_cdiv(m, t), (unsigned)omx_cdiv(n, t), (unsigned)batch);
    const bool one = k <= kMK;
#define OMX_SGEMM_LAUNCH(T, ONE)                                                                              \
    hipLaunchKernelGGL((k_sgemm_mfma<T, ONE>), grid, dim3(256), 0, st, m, n, k, alpha, A, lda, B, ldb, beta, C, \
                       ldc, sA, sB, sC)
    if (small) {
        if (one) OMX_SGEMM_LAUNCH(64, true);
        else OMX_SGEMM_LAUNCH(64, false);
    } else {
        if (one) OMX_SGEMM_LAUNCH(128, true);
        else OMX_SGEMM_LAUNCH(128, false);
    }
#undef OMX_SGEMM_LAUNCH
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

// msel: the row count the kernel is chosen for (default m).  A caller that
// skips rows known to be zero passes the full count, so that the rows it
// does compute round exactly as in the full product.
int sgemm(int64_t m, int64_t n, int64_t k, float alpha, const float *A, int64_t lda, const float *B,
          int64_t ldb, float beta, float *C, int64_t ldc, hipStream_t st, int64_t msel = -1) {
    if (m <= 0 || n <= 0) return OMX_OK;
    if (use_mfma(msel < 0 ? m : msel, n, k, 1, A, lda, B, ldb)) {
        return sgemm_mfma(m, n, k, alpha, A, lda, 0, B, ldb, 0, beta, C, ldc, 0, 1, st);
    }
    dim3 grid((unsigned)omx_cdiv(m, 64), (unsigned)omx_cdiv(n, 64));
    if (grid.y > 65535) return OMX_ERR_ARG;
    hipLaunchKernelGGL(k_sgemm_nn, grid, dim3(256), 0, st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                       (int64_t)0, (int64_t)0, (int64_t)0);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

// batch products C_p = beta C_p + alpha A_p B_p, X_p = X + p * sX
int sgemm_batched(int64_t m, int64_t n, int64_t k, float alpha, const float *A, int64_t lda, int64_t sA,
                  const float *B, int64_t ldb, int64_t sB, float beta, float *C, int64_t ldc, int64_t sC,
                  int64_t batch, hipStream_t st) {
    if (m <= 0 || n <= 0 || batch <= 0) return OMX_OK;
    dim3 grid((unsigned)omx_cdiv(m, 64), (unsigned)omx_cdiv(n, 64), (unsigned)batch);
    if (grid.y > 65535 || batch > 65535 || k % 16) return OMX_ERR_ARG;
    if (use_mfma(m, n, k, batch, A, lda, B, ldb) && (sA & 3) == 0 && (sB & 3) == 0) {
        return sgemm_mfma(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, st);
    }
    hipLaunchKernelGGL(k_sgemm_nn, grid, dim3(256), 0, st, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, sA, sB,
                       sC);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

// sgemm with n split into chunks of at most 65535*64 columns
int sgemm_big(int64_t m, int64_t n, int64_t k, float alpha, const float *A, int64_t lda, const float *B,
              int64_t ldb, float beta, float *C, int64_t ldc, hipStream_t st, int64_t msel = -1) {
    const int64_t nmax = (int64_t)65535 * 64;
    for (int64_t j = 0; j < n; j += nmax) {
        const int64_t nn = (n - j < nmax) ? n - j : nmax;
        int rc = sgemm(m, nn, k, alpha, A, lda, B + j * ldb, ldb, beta, C + j * ldc, ldc, st, msel);
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
        const int64_t rest = n - j - kNB;
        const int nrw = (int)omx_cdiv(rest, 64);
        const int prev = j > 0 ? 1 : 0;  // one more workgroup writes the previous step's factored D
        hipLaunchKernelGGL(k_getrf_step64, dim3((unsigned)(std::max(1, 2 * nrw) + prev)), dim3(256), 0, st, rest, nrw,
                           Ajj, lda, j, info, prev);
        OMX_LAUNCH_CHECK();
        if (rest <= 0) break;
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
            const unsigned g4 = (unsigned)omx_cdiv(m, 64);  // a lane quad per row
            for (int64_t J0 = 0; J0 < n; J0 += kOB) {
                const int64_t w = std::min(kOB, n - J0);
                for (int64_t j = J0; j < J0 + w; j += kNB) {
                    hipLaunchKernelGGL(k_trsm_rows_runn64, dim3(g4), dim3(256), 0, st, m, T + j + j * ldt, ldt,
                                       B + j * ldb, ldb);
                    OMX_LAUNCH_CHECK();
                    const int64_t inner = J0 + w - j - kNB;
                    if (inner > 0) RC(sgemm_big(m, inner, kNB, -1.0f, B + j * ldb, ldb, T + j + (j + kNB) * ldt, ldt, 1.0f,
                                                B + (j + kNB) * ldb, ldb, st));
                }
                const int64_t rest = n - J0 - w;
                if (rest > 0) RC(sgemm_big(m, rest, w, -1.0f, B + J0 * ldb, ldb, T + J0 + (J0 + w) * ldt, ldt, 1.0f,
                                           B + (J0 + w) * ldb, ldb, st));
            }
        } else {
            for (int64_t J1 = n; J1 > 0; J1 -= kOB) {
                const int64_t J0 = std::max<int64_t>(0, J1 - kOB);
                for (int64_t j = J1 - kNB; j >= J0; j -= kNB) {
                    hipLaunchKernelGGL(k_trsm_rows_rlnu64, dim3(g), dim3(256), 0, st, m, T + j + j * ldt, ldt,
                                       B + j * ldb, ldb);
                    OMX_LAUNCH_CHECK();
                    if (j > J0) RC(sgemm_big(m, j - J0, kNB, -1.0f, B + j * ldb, ldb, T + j + J0 * ldt, ldt, 1.0f,
                                             B + J0 * ldb, ldb, st));
                }
                if (J0 > 0) RC(sgemm_big(m, J0, J1 - J0, -1.0f, B + J0 * ldb, ldb, T + J0, ldt, 1.0f, B, ldb, st));
            }
        }
        return OMX_OK;
    }
    if (variant == OMX_TRSM_LLNU) {
        // inverse of an m x m unit lower triangle times B (m x n), m % 64 == 0
        if (m % kNB || ldt < m) return OMX_ERR_ARG;
        const int64_t nmax = (int64_t)65535 * 64;  // a lane quad per column
        for (int64_t I0 = 0; I0 < m; I0 += kOB) {
            const int64_t w = std::min(kOB, m - I0);
            for (int64_t i = I0; i < I0 + w; i += kNB) {
                for (int64_t c = 0; c < n; c += nmax) {
                    const int64_t nn = (n - c < nmax) ? n - c : nmax;
                    hipLaunchKernelGGL(k_trsm_cols_llnu64, dim3((unsigned)omx_cdiv(nn, 64)), dim3(256), 0, st, nn,
                                       T + i + i * ldt, ldt, B + i + c * ldb, ldb);
                    OMX_LAUNCH_CHECK();
                }
                const int64_t inner = I0 + w - i - kNB;
                if (inner > 0) RC(sgemm_big(inner, n, kNB, -1.0f, T + (i + kNB) + i * ldt, ldt, B + i, ldb, 1.0f,
                                            B + i + kNB, ldb, st));
            }
            const int64_t rest = m - I0 - w;
            if (rest > 0) RC(sgemm_big(rest, n, w, -1.0f, T + (I0 + w) + I0 * ldt, ldt, B + I0, ldb, 1.0f,
                                       B + I0 + w, ldb, st));
        }
        return OMX_OK;
    }
    return OMX_ERR_ARG;
}

int omx_trtri_lu_f32(int64_t n, const float *D, int64_t ldd, float *Xu, float *Xl, int64_t ldx, float *work,
                     void *stream) {
    if (n <= 0 || n % kNB || n > 65535 || !D || !Xu || !Xl || !work || ldd < n || ldx < n) return OMX_ERR_ARG;
    hipStream_t st = omx_stream(stream);
    const int nblk = (int)(n / kNB);
    hipLaunchKernelGGL(k_zero_offdiag, dim3((unsigned)n), dim3(256), 0, st, n, Xu, Xl, ldx);
    OMX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tri_inv64, dim3((unsigned)(2 * nblk)), dim3(256), 0, st, nblk, D, ldd, Xu, Xl, ldx);
    OMX_LAUNCH_CHECK();
    // Merge adjacent ranges of diagonal blocks level by level: ranges of s0
    // blocks pair up (the last, odd one waits for a later level).  For the
    // pair [R1 | R2] with inverses X1, X2 on the diagonal:
    //   U: X12 = -X1 (U12 X2)      L: X21 = -X2 (L21 X1)
    // with the inner product in work (n x n, leading dimension n, at the
    // position of the block it becomes).  Pairs of equal sizes run batched.
    std::vector<int64_t> sz(nblk, kNB);  // range sizes in rows, in order
    while (sz.size() > 1) {
        std::vector<int64_t> next;
        int64_t pos = 0;
        size_t i = 0;
        while (i + 1 < sz.size()) {
            // a batch: consecutive pairs with the same (s1, s2)
            const int64_t s1 = sz[i], s2 = sz[i + 1];
            int64_t cnt = 0;
            while (i + 1 < sz.size() && sz[i] == s1 && sz[i + 1] == s2) {
                ++cnt;
                i += 2;
            }
            const int64_t stride = s1 + s2;  // along the diagonal, per pair
            const int64_t a0 = pos, a1 = pos + s1;
            int rc;
            // U: work(a0.., a1..) = U12 X2;  Xu(a0.., a1..) = -X1 work
            rc = sgemm_batched(s1, s2, s2, 1.0f, D + a0 + a1 * ldd, ldd, stride * (ldd + 1), Xu + a1 + a1 * ldx, ldx,
                               stride * (ldx + 1), 0.0f, work + a0 + a1 * n, n, stride * (n + 1), cnt, st);
            if (rc) return rc;
            rc = sgemm_batched(s1, s2, s1, -1.0f, Xu + a0 + a0 * ldx, ldx, stride * (ldx + 1), work + a0 + a1 * n, n,
                               stride * (n + 1), 0.0f, Xu + a0 + a1 * ldx, ldx, stride * (ldx + 1), cnt, st);
            if (rc) return rc;
            // L: work(a1.., a0..) = L21 X1;  Xl(a1.., a0..) = -X2 work
            rc = sgemm_batched(s2, s1, s1, 1.0f, D + a1 + a0 * ldd, ldd, stride * (ldd + 1), Xl + a0 + a0 * ldx, ldx,
                               stride * (ldx + 1), 0.0f, work + a1 + a0 * n, n, stride * (n + 1), cnt, st);
            if (rc) return rc;
            rc = sgemm_batched(s2, s1, s2, -1.0f, Xl + a1 + a1 * ldx, ldx, stride * (ldx + 1), work + a1 + a0 * n, n,
                               stride * (n + 1), 0.0f, Xl + a1 + a0 * ldx, ldx, stride * (ldx + 1), cnt, st);
            if (rc) return rc;
            for (int64_t c = 0; c < cnt; ++c) next.push_back(stride);
            pos += cnt * stride;
        }
        if (i < sz.size()) next.push_back(sz[i]);
        sz.swap(next);
    }
    return OMX_OK;
}

int omx_trsm_identity_f32(int variant, int64_t n, const float *T, int64_t ldt, float *X, int64_t ldx,
                          void *stream) {
    if (n <= 0 || n % kNB || n > 65535 || !T || !X || ldt < n || ldx < n) return OMX_ERR_ARG;
    if (variant != OMX_TRSM_RUNN && variant != OMX_TRSM_RLNU) return OMX_ERR_ARG;
    hipStream_t st = omx_stream(stream);
    hipLaunchKernelGGL(k_identity, dim3((unsigned)omx_cdiv(n, 256), (unsigned)n), dim3(256), 0, st, n, X, ldx);
    OMX_LAUNCH_CHECK();
    // the blocking of omx_trsm_f32, the products restricted to the rows
    // that can be nonzero, the kernels chosen for all n rows: the results are
    // those of omx_trsm_f32 on the identity
    if (variant == OMX_TRSM_RUNN) {
        // block column j of I * U^-1 is zero below row j + 63
        for (int64_t J0 = 0; J0 < n; J0 += kOB) {
            const int64_t w = std::min(kOB, n - J0);
            for (int64_t j = J0; j < J0 + w; j += kNB) {
                const int64_t rows = j + kNB;
                hipLaunchKernelGGL(k_trsm_rows_runn64, dim3((unsigned)omx_cdiv(rows, 64)), dim3(256), 0, st, rows,
                                   T + j + j * ldt, ldt, X + j * ldx, ldx);
                OMX_LAUNCH_CHECK();
                const int64_t inner = J0 + w - j - kNB;
                if (inner > 0) RC(sgemm_big(rows, inner, kNB, -1.0f, X + j * ldx, ldx, T + j + (j + kNB) * ldt, ldt,
                                            1.0f, X + (j + kNB) * ldx, ldx, st, n));
            }
            const int64_t rest = n - J0 - w;
            if (rest > 0) RC(sgemm_big(J0 + w, rest, w, -1.0f, X + J0 * ldx, ldx, T + J0 + (J0 + w) * ldt, ldt, 1.0f,
                                       X + (J0 + w) * ldx, ldx, st, n));
        }
        return OMX_OK;
    }
    // block column j of I * L^-1 is zero above row j
    for (int64_t J1 = n; J1 > 0; J1 -= kOB) {
        const int64_t J0 = std::max<int64_t>(0, J1 - kOB);
        for (int64_t j = J1 - kNB; j >= J0; j -= kNB) {
            const int64_t rows = n - j;
            hipLaunchKernelGGL(k_trsm_rows_rlnu64, dim3((unsigned)omx_cdiv(rows, 256)), dim3(256), 0, st, rows,
                               T + j + j * ldt, ldt, X + j + j * ldx, ldx);
            OMX_LAUNCH_CHECK();
            if (j > J0) RC(sgemm_big(rows, j - J0, kNB, -1.0f, X + j + j * ldx, ldx, T + j + J0 * ldt, ldt, 1.0f,
                                     X + j + J0 * ldx, ldx, st, n));
        }
        if (J0 > 0) RC(sgemm_big(n - J0, J0, J1 - J0, -1.0f, X + J0 + J0 * ldx, ldx, T + J0, ldt, 1.0f, X + J0, ldx,
                                 st, n));
    }
    return OMX_OK;
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
