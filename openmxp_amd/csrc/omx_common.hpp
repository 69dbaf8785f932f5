// omx_common.hpp -- shared device/host helpers of the omxp HIP library.
//
// The 64-bit LCG here restates RandCoeff / RandStat of the reference
// (hpl_rand.hpp:12-78): x <- a*x + c mod 2^64 with a = 6364136223846793005,
// c = 1; element (i, j) of the order-n matrix is the state after i + n*j
// steps from initialize(seed) = a*seed + c; its value is
// double(int64(x)) * (2^-64 - 2^-117).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "omxp.h"

#define OMX_LCG_A 6364136223846793005ULL
#define OMX_LCG_C 1ULL

struct omx_map {
    uint64_t a, c;
};

// f o g (apply g first), hpl_rand.hpp:18-20
__host__ __device__ __forceinline__ omx_map omx_compose(omx_map f, omx_map g) {
    return omx_map{f.a * g.a, f.a * g.c + f.c};
}

// the map of `steps` LCG steps by binary powering (hpl_rand.hpp:31-48)
__host__ __device__ inline omx_map omx_jump(uint64_t steps) {
    omx_map base{OMX_LCG_A, OMX_LCG_C};
    omx_map r{1ULL, 0ULL};
    while (steps) {
        if (steps & 1ULL) r = omx_compose(r, base);
        base = omx_compose(base, base);
        steps >>= 1;
    }
    return r;
}

__host__ __device__ __forceinline__ uint64_t omx_apply(omx_map f, uint64_t x) {
    return f.a * x + f.c;
}

__host__ __device__ __forceinline__ uint64_t omx_lcg_init(uint64_t seed) {
    return OMX_LCG_A * seed + OMX_LCG_C;
}

// RandStat::operator double (hpl_rand.hpp:64-65).  int64 -> double is done
// as fma(double(hi), 2^32, double(lo)): both parts are exact and the fma
// rounds the exact sum once, which is the round-to-nearest-even conversion.
__device__ __forceinline__ double omx_lcg_value(uint64_t x) {
    const int32_t hi = (int32_t)(x >> 32);
    const uint32_t lo = (uint32_t)x;
    const double v = __builtin_fma((double)hi, 4294967296.0, (double)lo);
    return v * 0x1.fffffffffffffp-65;
}

__host__ inline double omx_lcg_value_host(uint64_t x) {
    return (double)(int64_t)x * 0x1.fffffffffffffp-65;
}

// stream handle of the C-ABI (void*) to hipStream_t
static inline hipStream_t omx_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

#define OMX_LAUNCH_CHECK()                                  \
    do {                                                    \
        hipError_t e_ = hipGetLastError();                  \
        if (e_ != hipSuccess) return OMX_ERR_HIP;           \
    } while (0)

#define OMX_HIP_CHECK(x)                                    \
    do {                                                    \
        hipError_t e_ = (x);                                \
        if (e_ != hipSuccess) return OMX_ERR_HIP;           \
    } while (0)

__host__ __device__ __forceinline__ int64_t omx_cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
#define omx_cdiv_dev omx_cdiv
