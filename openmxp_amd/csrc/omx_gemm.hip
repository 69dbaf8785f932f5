// omx_gemm.hip -- the Schur-complement GEMM on MFMA:
//   C (FP32, column-major, ldc) = beta*C + alpha * X * Y^T
// with X (m x k) and Y (n x k) packed FP16 or scaled FP8 e4m3 panels
// (omx_layout.hpp), accumulation in FP32.  The reference's trailing update is
// this with alpha = -1, beta = 1 (rocblas_gemm_ex f16 x f16 -> f32,
// panel_trf_gdirect.hpp:905-909); its -alt panel GEMMs use beta = 0
// (:158-162, :178-182).
//
// Workgroup tile 256 x 256, 4 waves in a 2 x 2 arrangement, each wave owns
// 128 x 128 of C as 4 x 4 accumulators of a 32 x 32 MFMA (256 FP32
// registers).  K advances in 32 KB stages per operand (64 FP16 or 128 FP8
// columns); two stages are double-buffered in LDS (128 KB).  Global -> LDS
// copies are direct (global_load_lds_dwordx4): the packed panel chunk is the
// LDS image, so no register staging is needed and every MFMA operand is one
// conflict-free ds_read_b128 per lane (two for FP8).
//
// The MFMA's A operand is the Y fragment and its B operand the X fragment,
// so in the 32 x 32 result lane l holds row i = l%32 of C and 16 columns:
// each store of a register writes 32 consecutive rows of one column.
#include "omx_common.hpp"
#include "omx_layout.hpp"

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef int intx8 __attribute__((ext_vector_type(8)));
typedef int intx4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void;

constexpr int kThreads = 256;
constexpr int kChunk = 32768;  // bytes per operand per stage
constexpr int kGroupM = 8;     // row strips per group of the tile order

// Workgroup -> (row tile, column tile).  Workgroups are dispatched
// round-robin over the 8 XCDs; consecutive logical ids are first gathered
// on one XCD (so one XCD's L2 sees a compact block of tiles), then ordered
// in groups of kGroupM row strips, column by column, so the workgroups in
// flight share X and Y strips.
__device__ __forceinline__ void tile_of(int bid, int tiles_m, int tiles_n, int &tm, int &tn) {
    const int total = tiles_m * tiles_n;
    int logical = bid;
    if ((total & 7) == 0) logical = (bid & 7) * (total >> 3) + (bid >> 3);
    const int gsize = kGroupM * tiles_n;
    const int g = logical / gsize;
    const int first = g * kGroupM;
    const int gm = min(tiles_m - first, kGroupM);
    const int in = logical - g * gsize;
    tm = first + in % gm;
    tn = in / gm;
}

__device__ __forceinline__ void wait_vmem() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Copy one 32 KB chunk global -> LDS; wave w moves bytes [8w KB, 8(w+1) KB).
__device__ __forceinline__ void load_chunk(const char *__restrict__ g, char *l, int wave, int lane) {
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int seg = wave * 8 + s;
        __builtin_amdgcn_global_load_lds(reinterpret_cast<const void *>(g + seg * 1024 + lane * 16),
                                         (lds_void *)(l + seg * 1024), 16, 0, 0);
    }
}

template <bool kBetaZero>
__device__ __forceinline__ void epilogue(floatx16 (&acc)[4][4], int64_t M, int64_t N, float alpha,
                                         float beta, float *__restrict__ C, int64_t ldc, int tm,
                                         int tn, int wm, int wn, int lane) {
    const int64_t ib = (int64_t)tm * OMX_STRIP + wm * 128 + (lane & 31);
    const int64_t jb = (int64_t)tn * OMX_STRIP + wn * 128 + 4 * (lane >> 5);
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
        const int64_t i = ib + mi * 32;
        if (i >= M) continue;
#pragma unroll
        for (int ni = 0; ni < 4; ++ni) {
            float *cp = C + i;
            float old[16];
            if (!kBetaZero) {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int64_t j = jb + ni * 32 + 8 * (r >> 2) + (r & 3);
                    old[r] = (j < N) ? cp[j * ldc] : 0.0f;
                }
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t j = jb + ni * 32 + 8 * (r >> 2) + (r & 3);
                if (j < N) {
                    const float v = acc[mi][ni][r];
                    cp[j * ldc] = kBetaZero ? alpha * v : beta * old[r] + alpha * v;
                }
            }
        }
    }
}

template <bool kBetaZero>
__global__ void __launch_bounds__(kThreads, 1)
k_gemm_f16(int64_t M, int64_t N, int64_t K, float alpha, float beta, const char *__restrict__ X,
           const char *__restrict__ Y, float *__restrict__ C, int64_t ldc, int tiles_m, int tiles_n) {
    __shared__ __attribute__((aligned(1024))) char smem[2][2][kChunk];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    int tm, tn;
    tile_of(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const int64_t nks = K / OMX_KSTAGE16;
    const char *xs = X + (int64_t)tm * nks * kChunk;
    const char *ys = Y + (int64_t)tn * nks * kChunk;

    floatx16 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = floatx16{};

    load_chunk(xs, smem[0][0], wave, lane);
    load_chunk(ys, smem[0][1], wave, lane);
    wait_vmem();
    __syncthreads();
    for (int64_t ks = 0; ks < nks; ++ks) {
        const int buf = (int)(ks & 1);
        if (ks + 1 < nks) {
            load_chunk(xs + (ks + 1) * kChunk, smem[buf ^ 1][0], wave, lane);
            load_chunk(ys + (ks + 1) * kChunk, smem[buf ^ 1][1], wave, lane);
        }
        const _Float16 *xl = reinterpret_cast<const _Float16 *>(smem[buf][0]);
        const _Float16 *yl = reinterpret_cast<const _Float16 *>(smem[buf][1]);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            half8 xf[4], yf[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                xf[q] = *reinterpret_cast<const half8 *>(xl + (((wm * 4 + q) * 4 + kk) * 512 + lane * 8));
                yf[q] = *reinterpret_cast<const half8 *>(yl + (((wn * 4 + q) * 4 + kk) * 512 + lane * 8));
            }
#pragma unroll
            for (int mi = 0; mi < 4; ++mi)
#pragma unroll
                for (int ni = 0; ni < 4; ++ni)
                    acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_f16(yf[ni], xf[mi], acc[mi][ni], 0, 0, 0);
        }
        wait_vmem();
        __syncthreads();
    }
    epilogue<kBetaZero>(acc, M, N, alpha, beta, C, ldc, tm, tn, wm, wn, lane);
}

template <bool kBetaZero>
__global__ void __launch_bounds__(kThreads, 1)
k_gemm_f8(int64_t M, int64_t N, int64_t K, float alpha, float beta, const char *__restrict__ X,
          const int32_t *__restrict__ xexp, const char *__restrict__ Y, const int32_t *__restrict__ yexp,
          float *__restrict__ C, int64_t ldc, int tiles_m, int tiles_n) {
    __shared__ __attribute__((aligned(1024))) char smem[2][2][kChunk];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    int tm, tn;
    tile_of(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const int64_t nks = K / OMX_KSTAGE8;
    const int64_t KT = K / OMX_SCALE_TILE;
    const char *xs = X + (int64_t)tm * nks * kChunk;
    const char *ys = Y + (int64_t)tn * nks * kChunk;
    const int32_t *xe = xexp + (int64_t)tm * KT;
    const int32_t *ye = yexp + (int64_t)tn * KT;

    floatx16 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = floatx16{};

    load_chunk(xs, smem[0][0], wave, lane);
    load_chunk(ys, smem[0][1], wave, lane);
    wait_vmem();
    __syncthreads();
    for (int64_t ks = 0; ks < nks; ++ks) {
        const int buf = (int)(ks & 1);
        if (ks + 1 < nks) {
            load_chunk(xs + (ks + 1) * kChunk, smem[buf ^ 1][0], wave, lane);
            load_chunk(ys + (ks + 1) * kChunk, smem[buf ^ 1][1], wave, lane);
        }
        // E8M0 scales: value = fp8 * 2^(s - 127) with s = 127 - e
        const int kt = (int)((ks * OMX_KSTAGE8) / OMX_SCALE_TILE);
        const int sx = 127 - xe[kt];
        const int sy = 127 - ye[kt];
        const char *xl = smem[buf][0];
        const char *yl = smem[buf][1];
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
            intx8 xf[4], yf[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int fx = ((wm * 4 + q) * 2 + kk) * 2048 + lane * 16;
                const int fy = ((wn * 4 + q) * 2 + kk) * 2048 + lane * 16;
                const intx4 xlo = *reinterpret_cast<const intx4 *>(xl + fx);
                const intx4 xhi = *reinterpret_cast<const intx4 *>(xl + fx + 1024);
                const intx4 ylo = *reinterpret_cast<const intx4 *>(yl + fy);
                const intx4 yhi = *reinterpret_cast<const intx4 *>(yl + fy + 1024);
                xf[q] = intx8{xlo[0], xlo[1], xlo[2], xlo[3], xhi[0], xhi[1], xhi[2], xhi[3]};
                yf[q] = intx8{ylo[0], ylo[1], ylo[2], ylo[3], yhi[0], yhi[1], yhi[2], yhi[3]};
            }
#pragma unroll
            for (int mi = 0; mi < 4; ++mi)
#pragma unroll
                for (int ni = 0; ni < 4; ++ni)
                    acc[mi][ni] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                        yf[ni], xf[mi], acc[mi][ni], 0, 0, 0, sy, 0, sx);
        }
        wait_vmem();
        __syncthreads();
    }
    epilogue<kBetaZero>(acc, M, N, alpha, beta, C, ldc, tm, tn, wm, wn, lane);
}

int gemm_args(int64_t m, int64_t n, int64_t k, int64_t kstage, const void *X, int64_t xr0,
              const void *Y, int64_t yr0, const float *C, int64_t ldc) {
    if (m < 0 || n < 0 || k <= 0 || k % kstage || !X || !Y || !C || ldc < m) return OMX_ERR_ARG;
    if (xr0 < 0 || yr0 < 0 || xr0 % OMX_STRIP || yr0 % OMX_STRIP) return OMX_ERR_ARG;
    const int64_t tiles = omx_cdiv(m, OMX_STRIP) * omx_cdiv(n, OMX_STRIP);
    if (tiles >= ((int64_t)1 << 31)) return OMX_ERR_ARG;
    return OMX_OK;
}

}  // namespace

extern "C" {

int omx_gemm_f16_packed(int64_t m, int64_t n, int64_t k, float alpha, const uint16_t *X,
                        int64_t x_row0, const uint16_t *Y, int64_t y_row0, float beta, float *C,
                        int64_t ldc, void *stream) {
    int rc = gemm_args(m, n, k, OMX_KSTAGE16, X, x_row0, Y, y_row0, C, ldc);
    if (rc) return rc;
    if (m == 0 || n == 0) return OMX_OK;
    const int tiles_m = (int)omx_cdiv(m, OMX_STRIP), tiles_n = (int)omx_cdiv(n, OMX_STRIP);
    const char *xp = reinterpret_cast<const char *>(X) + (x_row0 / OMX_STRIP) * OMX_STRIP * k * 2;
    const char *yp = reinterpret_cast<const char *>(Y) + (y_row0 / OMX_STRIP) * OMX_STRIP * k * 2;
    dim3 grid((unsigned)(tiles_m * tiles_n));
    if (beta == 0.0f)
        hipLaunchKernelGGL(k_gemm_f16<true>, grid, dim3(kThreads), 0, omx_stream(stream), m, n, k, alpha,
                           beta, xp, yp, C, ldc, tiles_m, tiles_n);
    else
        hipLaunchKernelGGL(k_gemm_f16<false>, grid, dim3(kThreads), 0, omx_stream(stream), m, n, k, alpha,
                           beta, xp, yp, C, ldc, tiles_m, tiles_n);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

int omx_gemm_f8_packed(int64_t m, int64_t n, int64_t k, float alpha, const uint8_t *X,
                       const int32_t *xexp, int64_t x_row0, const uint8_t *Y, const int32_t *yexp,
                       int64_t y_row0, float beta, float *C, int64_t ldc, void *stream) {
    int rc = gemm_args(m, n, k, OMX_SCALE_TILE, X, x_row0, Y, y_row0, C, ldc);
    if (rc) return rc;
    if (!xexp || !yexp) return OMX_ERR_ARG;
    if (m == 0 || n == 0) return OMX_OK;
    const int tiles_m = (int)omx_cdiv(m, OMX_STRIP), tiles_n = (int)omx_cdiv(n, OMX_STRIP);
    const int64_t KT = k / OMX_SCALE_TILE;
    const char *xp = reinterpret_cast<const char *>(X) + (x_row0 / OMX_STRIP) * OMX_STRIP * k;
    const char *yp = reinterpret_cast<const char *>(Y) + (y_row0 / OMX_STRIP) * OMX_STRIP * k;
    const int32_t *xe = xexp + (x_row0 / OMX_STRIP) * KT;
    const int32_t *ye = yexp + (y_row0 / OMX_STRIP) * KT;
    dim3 grid((unsigned)(tiles_m * tiles_n));
    if (beta == 0.0f)
        hipLaunchKernelGGL(k_gemm_f8<true>, grid, dim3(kThreads), 0, omx_stream(stream), m, n, k, alpha,
                           beta, xp, xe, yp, ye, C, ldc, tiles_m, tiles_n);
    else
        hipLaunchKernelGGL(k_gemm_f8<false>, grid, dim3(kThreads), 0, omx_stream(stream), m, n, k, alpha,
                           beta, xp, xe, yp, ye, C, ldc, tiles_m, tiles_n);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

}  // extern "C"
