// omx_pack.hip -- FP32 -> FP16 / scaled FP8 e4m3 panel packing into the GEMM
// layouts of omx_layout.hpp, plus the per-tile max-abs statistics used for
// the FP8 power-of-two scales and by the precision scheduler.
//
// A source element (r, k) of an R x K panel is src[r*rs + k*ls]: the left
// (column) panel A[k+1:, k] has rs = 1, ls = lda; the right (row) panel
// A[k, k+1:] read transposed has rs = lda, ls = 1 -- the two conversions of
// half_conversion_left / half_conversion_right_trans
// (fp16_gpu_kernels.cpp:11-36, :157-211).  Rows past R are packed as zeros.
#include "omx_common.hpp"
#include "omx_layout.hpp"

namespace {

constexpr int kPackThreads = 256;
constexpr int kTileK = 64;  // columns staged through LDS per workgroup

// mask on the packed coordinates (r, k) for the -alt 2 triangular copies
// (downCast_copy_upper / _lower, fp16_gpu_kernels.cpp:39-94)
__device__ __forceinline__ float apply_mask(float v, int64_t r, int64_t k, int mask) {
    switch (mask) {
        case OMX_MASK_K_LE_R: return (k <= r) ? v : 0.0f;
        case OMX_MASK_K_GT_R_UNIT: return (k > r) ? v : (k == r ? 1.0f : 0.0f);
        case OMX_MASK_R_GT_K_UNIT: return (r > k) ? v : (r == k ? 1.0f : 0.0f);
        default: return v;
    }
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// n consecutive-k values (n = 8 or 16, k from kb) of source row r (clamped
// to R - 1 so that every load is unconditional; the caller zeroes rows past
// R): the column panel (rs = 1) by n loads, which 32 lanes with consecutive
// rows turn into 128-byte runs, the row panel (ls = 1) by n/4 16-byte loads
// (the host checks 16-byte alignment of such a source).
template <int n>
__device__ __forceinline__ void load_run(const float *__restrict__ src, int64_t rs, int64_t ls, int64_t R, int64_t r,
                                         int64_t kb, float (&v)[n]) {
    const int64_t rc = r < R ? r : R - 1;
    if (rs == 1) {
#pragma unroll
        for (int e = 0; e < n; ++e) v[e] = src[rc + (kb + e) * ls];
    } else {
        const float *p = src + rc * rs + kb;
#pragma unroll
        for (int e = 0; e < n; e += 4) {
            const f32x4 w = *reinterpret_cast<const f32x4 *>(p + e);
            v[e] = w.x;
            v[e + 1] = w.y;
            v[e + 2] = w.z;
            v[e + 3] = w.w;
        }
    }
}

// One workgroup per (row strip, 64-column stage): 32 KB of packed output.
// Each thread writes 8 of the stage's 2048 16-byte units (8 consecutive k of
// one row); all 64 of its source values are loaded, straight from the panel,
// before the first is converted, so a workgroup's 64 KB arrive in one round
// trip.
__global__ void __launch_bounds__(kPackThreads)
k_pack_f16(int64_t R, int64_t K, const float *__restrict__ src, int64_t rs, int64_t ls,
           int mask, _Float16 *__restrict__ P, float *__restrict__ chunk_max,
           const int32_t *__restrict__ sexp, int64_t pchunks, int64_t pk0) {
    __shared__ float red[kPackThreads / 64];
    const int64_t ks = blockIdx.x;   // K-stage
    const int64_t rsb = blockIdx.y;  // row strip
    const int64_t r0 = rsb * OMX_STRIP, k0 = ks * kTileK;
    // the packed panel has pchunks 64-column chunks per strip; this panel's
    // columns start at chunk pk0
    _Float16 *out = P + (rsb * pchunks + pk0 + ks) * 16384;
    const int t = threadIdx.x;
    const float scale = sexp ? ldexpf(1.0f, sexp[rsb]) : 1.0f;
    float v[8][8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int oidx = t + kPackThreads * q;  // 16-byte unit, 0..2047
        const int fb = oidx >> 6, lane = oidx & 63;
        load_run<8>(src, rs, ls, R, r0 + (fb >> 2) * 32 + (lane & 31), k0 + (fb & 3) * 16 + (lane >> 5) * 8, v[q]);
    }
    float mx = 0.0f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int oidx = t + kPackThreads * q;
        const int fb = oidx >> 6, lane = oidx & 63;
        const int64_t r = r0 + (fb >> 2) * 32 + (lane & 31);
        const int64_t kb = k0 + (fb & 3) * 16 + (lane >> 5) * 8;
        typedef _Float16 half8 __attribute__((ext_vector_type(8)));
        half8 h;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float x = (r < R) ? apply_mask(v[q][e], r, kb + e, mask) : 0.0f;
            mx = fmaxf(mx, fabsf(x));
            h[e] = (_Float16)(x * scale);
        }
        *reinterpret_cast<half8 *>(out + oidx * 8) = h;
    }
    if (chunk_max) {
        for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
        if ((t & 63) == 0) red[t >> 6] = mx;
        __syncthreads();
        if (t == 0)
            chunk_max[rsb * (K >> 6) + ks] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    }
}

// Max |x| of each 256 x 256 (rows x K) scale tile.  Full tiles of a 16-byte
// aligned source: each thread takes 4 consecutive rows (column panel) or
// columns (row panel) of every fourth column (row), 16 loads in flight.
__global__ void __launch_bounds__(kPackThreads)
k_tile_maxabs(int64_t R, int64_t K, const float *__restrict__ src, int64_t rs, int64_t ls,
              float *__restrict__ tmax) {
    __shared__ float red[kPackThreads / 64];
    const int64_t kt = blockIdx.x, rsb = blockIdx.y;
    const int64_t r0 = rsb * OMX_STRIP, k0 = kt * OMX_SCALE_TILE;
    const int t = threadIdx.x;
    float mx = 0.0f;
    const int64_t step = (rs == 1) ? ls : rs;  // between the lines a thread walks
    const bool vec = (reinterpret_cast<uintptr_t>(src) & 15) == 0 && (step & 3) == 0 && r0 + OMX_STRIP <= R &&
                     k0 + OMX_SCALE_TILE <= K;
    if (vec) {
        const float *p = (rs == 1) ? src + r0 + 4 * (t & 63) + (k0 + (t >> 6)) * ls
                                   : src + (r0 + (t >> 6)) * rs + k0 + 4 * (t & 63);
#pragma unroll 16
        for (int j = 0; j < 64; ++j) {
            const f32x4 w = *reinterpret_cast<const f32x4 *>(p + (int64_t)(4 * j) * step);
            mx = fmaxf(mx, fmaxf(fmaxf(fabsf(w.x), fabsf(w.y)), fmaxf(fabsf(w.z), fabsf(w.w))));
        }
    } else if (rs == 1) {
        const int64_t r = r0 + t;
        if (r < R) {
#pragma unroll 8
            for (int64_t k = k0; k < k0 + OMX_SCALE_TILE && k < K; ++k) mx = fmaxf(mx, fabsf(src[r + k * ls]));
        }
    } else {
        const int64_t k = k0 + t;
#pragma unroll 8
        for (int q = 0; q < OMX_STRIP; ++q) {
            const int64_t r = r0 + q;
            if (r < R && k < K) mx = fmaxf(mx, fabsf(src[r * rs + k]));
        }
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if ((t & 63) == 0) red[t >> 6] = mx;
    __syncthreads();
    if (t == 0) tmax[rsb * omx_cdiv_dev(K, OMX_SCALE_TILE) + kt] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

// largest e with maxabs * 2^e <= 448, clamped to [-100, 100]; 0 for zero
__device__ __forceinline__ int pow2_exp_448(float maxabs) {
    if (!(maxabs > 0.0f) || !(maxabs < __builtin_inff())) return 0;
    int em;
    const float fm = frexpf(maxabs, &em);  // maxabs = fm * 2^em, fm in [0.5, 1)
    int e = (fm <= 0.875f) ? (9 - em) : (8 - em);  // 448 = 0.875 * 2^9
    return e > 100 ? 100 : (e < -100 ? -100 : e);
}

// exps[strip * ekt + ek0 + t] from tmax[strip * KT + t] (the exponent array
// may span a wider packed panel than the tile maxima)
__global__ void k_tile_exps(int64_t n, int64_t KT, const float *__restrict__ tmax, int headroom,
                            int32_t *__restrict__ exps, int64_t ekt, int64_t ek0) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const int e = pow2_exp_448(tmax[i]);
        exps[(i / KT) * ekt + ek0 + i % KT] = (tmax[i] > 0.0f) ? (e - headroom < -100 ? -100 : e - headroom) : 0;
    }
}

// Per-strip FP16 scale: the largest e with max|strip| * 2^e <= 2^15 (half the
// FP16 range, clear of overflow), clamped to [-100, 100]; 0 for a zero strip.
__global__ void k_strip_exps16(int64_t nstrips, int64_t KT, const float *__restrict__ tmax,
                               int32_t *__restrict__ sexp) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nstrips) return;
    float m = 0.0f;
    for (int64_t q = 0; q < KT; ++q) m = fmaxf(m, tmax[s * KT + q]);
    int e = 0;
    if (m > 0.0f && m < __builtin_inff()) {
        int em;
        frexpf(m, &em);  // m in [2^(em-1), 2^em)
        e = 15 - em;     // m * 2^e < 2^15
        e = e > 100 ? 100 : (e < -100 ? -100 : e);
    }
    sexp[s] = e;
}

// FP8 pack: one workgroup per (row strip, 64-column half of a 128 stage);
// each thread writes 4 16-byte units (16 consecutive k of one row), its 64
// source values loaded straight from the panel first.
__global__ void __launch_bounds__(kPackThreads)
k_pack_f8(int64_t R, int64_t K, const float *__restrict__ src, int64_t rs, int64_t ls,
          const int32_t *__restrict__ exps, uint8_t *__restrict__ P, int64_t pchunks, int64_t pk0,
          int64_t ekt, int64_t ek0) {
    const int64_t k64 = blockIdx.x;  // 64-column block
    const int64_t rsb = blockIdx.y;
    const int64_t r0 = rsb * OMX_STRIP, k0 = k64 * kTileK;
    const float scale = ldexpf(1.0f, exps[rsb * ekt + ek0 + k0 / OMX_SCALE_TILE]);
    const int64_t ks = k64 >> 1;
    const int kk = (int)(k64 & 1);
    // pchunks 128-column chunks per strip; this panel starts at chunk pk0
    uint8_t *chunk = P + (rsb * pchunks + pk0 + ks) * 32768;
    const int t = threadIdx.x;
    float v[4][16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int oidx = t + kPackThreads * q;  // 0..1023: (rr, h, lane)
        const int rr = oidx >> 7, h = (oidx >> 6) & 1, lane = oidx & 63;
        load_run<16>(src, rs, ls, R, r0 + rr * 32 + (lane & 31), k0 + (lane >> 5) * 32 + h * 16, v[q]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int oidx = t + kPackThreads * q;
        const int rr = oidx >> 7, h = (oidx >> 6) & 1, lane = oidx & 63;
        const bool in = r0 + rr * 32 + (lane & 31) < R;
        uint32_t w[4];
#pragma unroll
        for (int e4 = 0; e4 < 4; ++e4) {
            const float *x = &v[q][4 * e4];
            int pk = 0;
            pk = __builtin_amdgcn_cvt_pk_fp8_f32(in ? x[0] * scale : 0.0f, in ? x[1] * scale : 0.0f, pk, false);
            pk = __builtin_amdgcn_cvt_pk_fp8_f32(in ? x[2] * scale : 0.0f, in ? x[3] * scale : 0.0f, pk, true);
            w[e4] = (uint32_t)pk;
        }
        uint8_t *dst = chunk + ((rr * 2 + kk) * 2048) + h * 1024 + lane * 16;
        *reinterpret_cast<uint4 *>(dst) = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

}  // namespace

extern "C" {

size_t omx_panel_bytes(int64_t rows, int64_t k, int fp8) {
    if (rows <= 0 || k <= 0) return 0;
    return (size_t)omx_strips(rows) * OMX_STRIP * (size_t)k * (fp8 ? 1 : 2);
}

size_t omx_panel_exps_count(int64_t rows, int64_t k) {
    if (rows <= 0 || k <= 0) return 0;
    return (size_t)omx_strips(rows) * (size_t)omx_cdiv(k, OMX_SCALE_TILE);
}

int omx_pack_f16(int64_t rows, int64_t k, const float *src, int64_t rs, int64_t ls, int mask,
                 uint16_t *P, float *chunk_max, void *stream) {
    if (rows <= 0 || k <= 0 || k % OMX_KSTAGE16 || !src || !P || mask < 0 || mask > 3) return OMX_ERR_ARG;
    if (!((rs == 1 && ls >= rows) || (ls == 1 && rs >= k))) return OMX_ERR_ARG;
    if (ls == 1 && (rs % 4 || (reinterpret_cast<uintptr_t>(src) & 15))) return OMX_ERR_ARG;
    dim3 grid((unsigned)(k / kTileK), (unsigned)omx_strips(rows));
    hipLaunchKernelGGL(k_pack_f16, grid, dim3(kPackThreads), 0, omx_stream(stream), rows, k, src, rs,
                       ls, mask, reinterpret_cast<_Float16 *>(P), chunk_max, nullptr, k / 64, (int64_t)0);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

int omx_pack_f16_scaled(int64_t rows, int64_t k, const float *src, int64_t rs, int64_t ls, int have_max,
                        uint16_t *P, int32_t *sexp, float *tile_max, void *stream) {
    if (rows <= 0 || k <= 0 || k % OMX_SCALE_TILE || !src || !P || !sexp || !tile_max) return OMX_ERR_ARG;
    if (!((rs == 1 && ls >= rows) || (ls == 1 && rs >= k))) return OMX_ERR_ARG;
    if (ls == 1 && (rs % 4 || (reinterpret_cast<uintptr_t>(src) & 15))) return OMX_ERR_ARG;
    hipStream_t st = omx_stream(stream);
    const int64_t strips = omx_strips(rows), KT = omx_cdiv(k, OMX_SCALE_TILE);
    if (!have_max) {
        hipLaunchKernelGGL(k_tile_maxabs, dim3((unsigned)KT, (unsigned)strips), dim3(kPackThreads), 0, st,
                           rows, k, src, rs, ls, tile_max);
        OMX_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_strip_exps16, dim3((unsigned)omx_cdiv(strips, 256)), dim3(256), 0, st, strips, KT, tile_max,
                       sexp);
    OMX_LAUNCH_CHECK();
    dim3 grid((unsigned)(k / kTileK), (unsigned)strips);
    hipLaunchKernelGGL(k_pack_f16, grid, dim3(kPackThreads), 0, st, rows, k, src, rs, ls, (int)OMX_MASK_NONE,
                       reinterpret_cast<_Float16 *>(P), nullptr, sexp, k / 64, (int64_t)0);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

int omx_pack_f8_window(int64_t rows, int64_t k, const float *src, int64_t rs, int64_t ls, int headroom, int have_max,
                       uint8_t *P, int32_t *exps, int64_t ldk, int64_t k0, float *tile_max, void *stream) {
    if (rows <= 0 || k <= 0 || k % OMX_SCALE_TILE || !src || !P || !exps || !tile_max) return OMX_ERR_ARG;
    if (ldk <= 0 || ldk % OMX_SCALE_TILE || k0 < 0 || k0 % OMX_SCALE_TILE || k0 + k > ldk) return OMX_ERR_ARG;
    if (headroom < 0 || headroom > 3) return OMX_ERR_ARG;
    if (!((rs == 1 && ls >= rows) || (ls == 1 && rs >= k))) return OMX_ERR_ARG;
    if (ls == 1 && (rs % 4 || (reinterpret_cast<uintptr_t>(src) & 15))) return OMX_ERR_ARG;
    hipStream_t st = omx_stream(stream);
    const int64_t strips = omx_strips(rows), KT = omx_cdiv(k, OMX_SCALE_TILE);
    if (!have_max) {
        hipLaunchKernelGGL(k_tile_maxabs, dim3((unsigned)KT, (unsigned)strips), dim3(kPackThreads), 0, st,
                           rows, k, src, rs, ls, tile_max);
        OMX_LAUNCH_CHECK();
    }
    const int64_t nt = strips * KT;
    const int64_t ekt = ldk / OMX_SCALE_TILE, ek0 = k0 / OMX_SCALE_TILE;
    hipLaunchKernelGGL(k_tile_exps, dim3((unsigned)omx_cdiv(nt, 256)), dim3(256), 0, st, nt, KT, tile_max, headroom,
                       exps, ekt, ek0);
    OMX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pack_f8, dim3((unsigned)(k / kTileK), (unsigned)strips), dim3(kPackThreads), 0,
                       st, rows, k, src, rs, ls, exps, P, ldk / OMX_KSTAGE8, k0 / OMX_KSTAGE8, ekt, ek0);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

int omx_pack_f8_ex(int64_t rows, int64_t k, const float *src, int64_t rs, int64_t ls, int headroom, int have_max,
                   uint8_t *P, int32_t *exps, float *tile_max, void *stream) {
    return omx_pack_f8_window(rows, k, src, rs, ls, headroom, have_max, P, exps, k, 0, tile_max, stream);
}

int omx_pack_f16_window(int64_t rows, int64_t k, const float *src, int64_t rs, int64_t ls, int have_max,
                        uint16_t *P, int32_t *sexp, float *tile_max, int64_t ldk, int64_t k0, void *stream) {
    if (rows <= 0 || k <= 0 || k % OMX_KSTAGE16 || !src || !P) return OMX_ERR_ARG;
    if (ldk <= 0 || ldk % OMX_KSTAGE16 || k0 < 0 || k0 % OMX_KSTAGE16 || k0 + k > ldk) return OMX_ERR_ARG;
    if (sexp && (!tile_max || k % OMX_SCALE_TILE)) return OMX_ERR_ARG;
    if (!((rs == 1 && ls >= rows) || (ls == 1 && rs >= k))) return OMX_ERR_ARG;
    if (ls == 1 && (rs % 4 || (reinterpret_cast<uintptr_t>(src) & 15))) return OMX_ERR_ARG;
    hipStream_t st = omx_stream(stream);
    const int64_t strips = omx_strips(rows);
    if (sexp) {
        const int64_t KT = omx_cdiv(k, OMX_SCALE_TILE);
        if (!have_max) {
            hipLaunchKernelGGL(k_tile_maxabs, dim3((unsigned)KT, (unsigned)strips), dim3(kPackThreads), 0, st,
                               rows, k, src, rs, ls, tile_max);
            OMX_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(k_strip_exps16, dim3((unsigned)omx_cdiv(strips, 256)), dim3(256), 0, st, strips, KT,
                           tile_max, sexp);
        OMX_LAUNCH_CHECK();
    }
    dim3 grid((unsigned)(k / kTileK), (unsigned)strips);
    hipLaunchKernelGGL(k_pack_f16, grid, dim3(kPackThreads), 0, st, rows, k, src, rs, ls, (int)OMX_MASK_NONE,
                       reinterpret_cast<_Float16 *>(P), nullptr, sexp, ldk / 64, k0 / 64);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

int omx_pack_f8(int64_t rows, int64_t k, const float *src, int64_t rs, int64_t ls, uint8_t *P,
                int32_t *exps, float *tile_max, void *stream) {
    return omx_pack_f8_ex(rows, k, src, rs, ls, 0, 0, P, exps, tile_max, stream);
}

int omx_panel_tile_maxabs(int64_t rows, int64_t k, const float *src, int64_t rs, int64_t ls, float *tile_max,
                          void *stream) {
    if (rows <= 0 || k <= 0 || k % OMX_SCALE_TILE || !src || !tile_max) return OMX_ERR_ARG;
    if (!((rs == 1 && ls >= rows) || (ls == 1 && rs >= k))) return OMX_ERR_ARG;
    const int64_t strips = omx_strips(rows), KT = omx_cdiv(k, OMX_SCALE_TILE);
    hipLaunchKernelGGL(k_tile_maxabs, dim3((unsigned)KT, (unsigned)strips), dim3(kPackThreads), 0,
                       omx_stream(stream), rows, k, src, rs, ls, tile_max);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

}  // extern "C"
