// omx_layout.hpp -- the packed panel layouts read by the MFMA Schur GEMM.
//
// A panel is an R x K operand (R rows of the matrix, K = the block size b).
// Rows are grouped in strips of 256; every (strip, K-stage) pair is one
// contiguous 32 KB chunk, which is exactly one LDS stage of the GEMM, so a
// workgroup streams its operands with linear 16-byte loads.
//
// FP16 (K-stage 64): a chunk holds 32 fragment blocks of 1 KB, block
// index (row%256)/32 * 4 + (k%64)/16.  Inside a block, lane l = row%32 +
// 32*((k%16)/8) owns 8 consecutive halves (k%8), i.e. exactly the A/B
// operand of one v_mfma_f32_32x32x16_f16 for that lane.
//
// FP8 e4m3 (K-stage 128): a chunk holds 16 fragment blocks of 2 KB, block
// index (row%256)/32 * 2 + (k%128)/64.  Lane l = row%32 + 32*((k%64)/32)
// owns 32 bytes, stored as two 1 KB halves ((k%32)/16) of 16 bytes per lane,
// i.e. the operand of one v_mfma_scale_f32_32x32x64_f8f6f4.  Each 256 x 256
// (rows x K) tile has a power-of-two exponent e: value = fp8 * 2^-e; the
// exponents are stored strip-major, exps[(row/256) * ceil(K/256) + k/256].
#pragma once

#include <stdint.h>

#define OMX_STRIP 256
#define OMX_KSTAGE16 64
#define OMX_KSTAGE8 128
#define OMX_SCALE_TILE 256

__host__ __device__ __forceinline__ int64_t omx_pk16_offset(int64_t r, int64_t k, int64_t K) {
    return ((r >> 8) * (K >> 6) + (k >> 6)) * 16384 +
           ((((r & 255) >> 5) << 2) + ((k & 63) >> 4)) * 512 +
           (((r & 31) + 32 * ((k & 15) >> 3)) << 3) + (k & 7);
}

__host__ __device__ __forceinline__ int64_t omx_pk8_offset(int64_t r, int64_t k, int64_t K) {
    return ((r >> 8) * (K >> 7) + (k >> 7)) * 32768 +
           ((((r & 255) >> 5) << 1) + ((k & 127) >> 6)) * 2048 + ((k & 31) >> 4) * 1024 +
           (((r & 31) + 32 * ((k & 63) >> 5)) << 4) + (k & 15);
}

__host__ __device__ __forceinline__ int64_t omx_strips(int64_t rows) { return (rows + 255) >> 8; }
