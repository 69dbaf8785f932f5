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
// registers).  K advances in LDS stages of 16 KB per operand (32 FP16 or 64
// FP8 columns of the 256-row strips); OMX_GEMM_NBUF stages form a ring in LDS
// (128 KB for 4) with all but one in flight.  Global -> LDS copies are direct
// (global_load_lds_dwordx4): the packed panel chunk is the LDS image, so no
// register staging is needed and every MFMA operand is one conflict-free
// ds_read_b128 per lane (two for FP8).
//
// The MFMA's A operand is the Y fragment and its B operand the X fragment,
// so in the 32 x 32 result lane l holds row i = l%32 of C and 16 columns:
// each store of a register writes 32 consecutive rows of one column.
#include <stdlib.h>
#include <string.h>

#include <type_traits>

#include "omx_common.hpp"
#include "omx_layout.hpp"

#ifndef OMX_GEMM_NBUF
#define OMX_GEMM_NBUF 4
#endif

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
// round-robin over the 8 XCDs (workgroup id % 8); the ids are first remapped
// so that each XCD owns a contiguous range of logical tiles (its L2 then
// sees a compact block of the tile space), and logical tiles are ordered in
// groups of kGroupM row strips, column by column, so the workgroups in
// flight on one XCD share X and Y strips.
__device__ __forceinline__ void tile_of(int bid, int tiles_m, int tiles_n, int &tm, int &tn) {
    const int total = tiles_m * tiles_n;
    const int xcd = bid & 7, idx = bid >> 3;
    const int q = total >> 3, r = total & 7;
    const int logical = (xcd < r) ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
    const int gsize = kGroupM * tiles_n;
    const int g = logical / gsize;
    const int first = g * kGroupM;
    const int gm = min(tiles_m - first, kGroupM);
    const int in = logical - g * gsize;
    // uniform values: the divisions above go through VALU reciprocals, and
    // without the readfirstlane every buffer access whose descriptor derives
    // from the tile would be wrapped in a waterfall loop
    tm = __builtin_amdgcn_readfirstlane(first + in % gm);
    tn = __builtin_amdgcn_readfirstlane(in / gm);
}

// s_waitcnt vmcnt(8 * stages): wait until at most `stages` later stages
// (8 loads per wave each) are still in flight.  gfx9 encoding of the
// immediate: vmcnt[3:0] | expcnt[6:4] | lgkmcnt[11:8] | vmcnt[5:4] << 14;
// expcnt and lgkmcnt are left at their maxima (no wait).
__device__ __forceinline__ void wait_stages(int stages) {
    switch (stages) {
        case 0: __builtin_amdgcn_s_waitcnt(0x0F70); break;   // vmcnt(0)
        case 1: __builtin_amdgcn_s_waitcnt(0x0F78); break;   // vmcnt(8)
        case 2: __builtin_amdgcn_s_waitcnt(0x4F70); break;   // vmcnt(16)
        default: __builtin_amdgcn_s_waitcnt(0x4F78); break;  // vmcnt(24)
    }
}

// The stage barrier without __syncthreads()'s release/acquire fences: those
// make the compiler drain every outstanding load (vmcnt(0)) before the
// barrier, including the LDS-DMA copies of later stages.  The explicit vmcnt
// wait before it orders the copies the next stage reads, and the lgkmcnt
// wait the compiler puts before each MFMA completes the reads of a ring slot
// before any wave passes the barrier after which that slot is refilled.  The
// memory clobber keeps the compiler from moving LDS accesses across it.
// Used by both 256 x 256 kernels: +3 % on the FP16 trailing update; in the
// FP8 kernel, whose stages come in pairs, __syncthreads() made the second
// stage of each pair wait for every copy in flight (the refill issued eight
// MFMAs earlier included), a prefetch distance of one stage: +6 % (k = 10240)
// to +10 % (k = 5120) without it.
__device__ __forceinline__ void stage_barrier() { asm volatile("s_barrier" ::: "memory"); }

__device__ __forceinline__ void gl_lds16(const char *g, char *l) {
    __builtin_amdgcn_global_load_lds(reinterpret_cast<const void *>(g), (lds_void *)l, 16, 0, 0);
}

// C update of one wave's 128 x 128 block.  acc[mi][ni][r] is
//   C(i0 + 32 mi + lane%32, j0 + 32 ni + 8 (r/4) + 4 (lane/32) + r%4),
// so each lane owns 16 columns of one row per 32 x 32 block and every load or
// store instruction covers 32 consecutive rows of two columns.  Interior
// tiles go through buffer instructions, one resource per 32-column block
// (wave-uniform 64-bit base in SGPRs, 16 per-lane 32-bit column offsets, the
// row block in the immediate), pipelined over ni: the 64 old values of
// column block ni + 1 are in flight while column block ni is combined and
// stored.
// Cache policy of the C loads and stores (buffer aux operand): NT (bit 1),
// non-temporal.  Each C element is read and written once per launch, while
// the packed panels are re-read by many workgroups through L2; streaming C
// keeps its lines from displacing panel lines (FP8 trailing update at
// 61440 x 61440 x 2560: 1404 -> 1548 TFLOP/s, FP16 unchanged).
constexpr int kCPolicy = 2;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t colblock_rsrc(float *p, int64_t ldc) {
    return __builtin_amdgcn_make_buffer_rsrc(p, (short)0, (int)(ldc * 32 * 4), 0x00020000);
}

// mx (beta = 0): max |value stored| by this lane, folded into what it holds.
template <bool kBetaZero>
__device__ __forceinline__ void epilogue(floatx16 (&acc)[4][4], int64_t M, int64_t N, float alpha,
                                         float beta, float *__restrict__ C, int64_t ldc, int tm,
                                         int tn, int wm, int wn, int lane, float &mx) {
    const int64_t i0 = (int64_t)tm * OMX_STRIP + wm * 128;
    const int64_t j0 = (int64_t)tn * OMX_STRIP + wn * 128;
    if ((int64_t)(tm + 1) * OMX_STRIP <= M && (int64_t)(tn + 1) * OMX_STRIP <= N &&
        ldc * 32 * 4 < ((int64_t)1 << 31)) {
        int voff[16];
#pragma unroll
        for (int r = 0; r < 16; ++r)
            voff[r] = (int)(((8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)) * ldc + (lane & 31)) * 4);
        if (kBetaZero) {
#pragma unroll
            for (int ni = 0; ni < 4; ++ni) {
                const __amdgpu_buffer_rsrc_t rs = colblock_rsrc(C + (j0 + 32 * ni) * ldc + i0, ldc);
#pragma unroll
                for (int mi = 0; mi < 4; ++mi)
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const float v = alpha * acc[mi][ni][r];
                        mx = fmaxf(mx, fabsf(v));
                        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), rs, voff[r] + mi * 128,
                                                              0, kCPolicy);
                    }
            }
            return;
        }
        // old values in units of two row blocks of one column block (32
        // registers), the next unit's loads in flight while one is stored
        float old[2][2][16];
        auto fetch = [&](float(&o)[2][16], int u) {
            const int ni = u >> 1, m0 = 2 * (u & 1);
            const __amdgpu_buffer_rsrc_t rs = colblock_rsrc(C + (j0 + 32 * ni) * ldc + i0, ldc);
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    o[h][r] = __builtin_bit_cast(
                        float, __builtin_amdgcn_raw_buffer_load_b32(rs, voff[r] + (m0 + h) * 128, 0, kCPolicy));
        };
        fetch(old[0], 0);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (u + 1 < 8) fetch(old[(u + 1) & 1], u + 1);
            const int ni = u >> 1, m0 = 2 * (u & 1);
            const __amdgpu_buffer_rsrc_t rs = colblock_rsrc(C + (j0 + 32 * ni) * ldc + i0, ldc);
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    __builtin_amdgcn_raw_buffer_store_b32(
                        __builtin_bit_cast(uint32_t, beta * old[u & 1][h][r] + alpha * acc[m0 + h][ni][r]), rs,
                        voff[r] + (m0 + h) * 128, 0, kCPolicy);
        }
        return;
    }
    // edge tiles (and ldc too large for 32-bit column offsets): checked
    const int64_t ib = i0 + (lane & 31);
    const int64_t jb = j0 + 4 * (lane >> 5);
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
                    if (kBetaZero) {
                        const float w = alpha * v;
                        mx = fmaxf(mx, fabsf(w));
                        cp[j * ldc] = w;
                    } else {
                        cp[j * ldc] = beta * old[r] + alpha * v;
                    }
                }
            }
        }
    }
}

// The beta != 0 epilogue of an interior tile through the (then idle) LDS
// ring: each wave lands its C block in its own 32 KB of the ring by LDS-DMA,
// 16 KB (one 32-column block) per batch, two batches in flight — no staging
// registers, one 16-byte copy per lane per two columns — and reads the old
// values back from LDS in the accumulator layout (four row blocks per
// column block).  The VMEM order per wave is D0 D1 | D2 S0 | D3 S1 | S2 |
// S3 (D: a batch's 16 copies, S: its 64 stores), and vector-memory
// operations complete in issue order, so "batch ni has landed" is a
// vmcnt wait: 16 after D0, then at most 63 outstanding (the copies of the
// batch after next and part of the stores) for the later batches.  The
// caller has passed a barrier after the last stage (no wave still reads
// the ring); C 16-byte aligned, ldc % 4 == 0.  kRows row blocks of a batch
// are read back and stored at a time: 4 in the FP8 kernel; 2 in the FP16
// one, whose main loop leaves fewer registers free (there the refill comes
// after the first 32 stores of its batch, so the vmcnt(63) waits are
// stricter than needed, never weaker).
__device__ __forceinline__ bool c_lds_ok(int64_t M, int64_t N, const float *C, int64_t ldc, int tm, int tn) {
    return (int64_t)(tm + 1) * OMX_STRIP <= M && (int64_t)(tn + 1) * OMX_STRIP <= N &&
           ldc * 32 * 4 < ((int64_t)1 << 31) && ldc % 4 == 0 && (reinterpret_cast<uintptr_t>(C) & 15) == 0;
}

template <int kRows>
__device__ __forceinline__ void epilogue_lds(floatx16 (&acc)[4][4], float alpha, float beta,
                                             float *__restrict__ C, int64_t ldc, int64_t i0, int64_t j0,
                                             int lane, int wave, char *ring) {
    char *reg = ring + wave * 32768;
    int voff[16];
#pragma unroll
    for (int r = 0; r < 16; ++r)
        voff[r] = (int)(((8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)) * ldc + (lane & 31)) * 4);
    // batch ni into half h: copy t brings columns 2t (lanes 0-31) and 2t + 1
    // (lanes 32-63) of the block, 128 rows each: LDS column c at c * 512
    auto dma = [&](int ni, int h) {
        const char *g = reinterpret_cast<const char *>(C + (j0 + 32 * ni + (lane >> 5)) * ldc + i0 + 4 * (lane & 31));
        char *l = reg + h * 16384;
#pragma unroll
        for (int t = 0; t < 16; ++t) gl_lds16(g + (int64_t)2 * t * ldc * 4, l + t * 1024);
    };
    // the half of batch ni is refilled with batch ni + 2 once its reads have
    // completed
    auto refill = [&](int ni, int h) {
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
        __builtin_amdgcn_sched_barrier(0);
        dma(ni + 2, h);
    };
    dma(0, 0);
    dma(1, 1);
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
        const int h = ni & 1;
        __builtin_amdgcn_sched_barrier(0);
        if (ni == 0)
            __builtin_amdgcn_s_waitcnt(0x4F70);  // vmcnt(16)
        else
            __builtin_amdgcn_s_waitcnt(0xCF7F);  // vmcnt(63)
        __builtin_amdgcn_sched_barrier(0);
        const char *l = reg + h * 16384;
        if constexpr (kRows == 4) {
            float old[4][16];
#pragma unroll
            for (int mi = 0; mi < 4; ++mi)
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    old[mi][r] = *reinterpret_cast<const float *>(
                        l + (8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)) * 512 + (32 * mi + (lane & 31)) * 4);
            if (ni + 2 < 4) refill(ni, h);
            const __amdgpu_buffer_rsrc_t rs = colblock_rsrc(C + (j0 + 32 * ni) * ldc + i0, ldc);
#pragma unroll
            for (int mi = 0; mi < 4; ++mi)
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    __builtin_amdgcn_raw_buffer_store_b32(
                        __builtin_bit_cast(uint32_t, beta * old[mi][r] + alpha * acc[mi][ni][r]), rs,
                        voff[r] + mi * 128, 0, kCPolicy);
        } else {
            const __amdgpu_buffer_rsrc_t rs = colblock_rsrc(C + (j0 + 32 * ni) * ldc + i0, ldc);
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                float old[2][16];
#pragma unroll
                for (int q = 0; q < 2; ++q)
#pragma unroll
                    for (int r = 0; r < 16; ++r)
                        old[q][r] = *reinterpret_cast<const float *>(
                            l + (8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)) * 512 + (32 * (2 * half + q) + (lane & 31)) * 4);
                if (half == 1 && ni + 2 < 4) refill(ni, h);
#pragma unroll
                for (int q = 0; q < 2; ++q)
#pragma unroll
                    for (int r = 0; r < 16; ++r)
                        __builtin_amdgcn_raw_buffer_store_b32(
                            __builtin_bit_cast(uint32_t, beta * old[q][r] + alpha * acc[2 * half + q][ni][r]), rs,
                            voff[r] + (2 * half + q) * 128, 0, kCPolicy);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// FP16: LDS stages of 32 columns (16 KB per operand), kNbuf of them in a
// ring, kNbuf-1 in flight.  Stage s of a 256-row strip is half (s & 1) of
// packed chunk s/2: fragment blocks rr*4 + 2*(s&1) + {0,1}, copied to the
// LDS stage as 16 blocks of 1 KB in [rr][k-step] order.  Per stage each wave
// runs two k-steps of 16 MFMAs; between them it waits for the next stage,
// passes the single barrier of the stage, refills the ring slot freed by the
// previous stage and reads the next stage's first fragments, which then
// arrive while the second 16 MFMAs run.
// ---------------------------------------------------------------------------
template <bool kBetaZero, int kNbuf, bool kTmax = false>
__global__ void __launch_bounds__(kThreads, 1)
k_gemm_f16(int64_t M, int64_t N, int64_t K, float alpha, float beta, const char *__restrict__ X,
           const char *__restrict__ Y, float *__restrict__ C, int64_t ldc, int tiles_m, int tiles_n,
           const int32_t *__restrict__ xsexp, const int32_t *__restrict__ ysexp, int64_t xstrip,
           int64_t ystrip, int ktri, float *__restrict__ tmax, int tmax_tr) {
    __shared__ __attribute__((aligned(1024))) char smem[kNbuf][2][16384];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: scalar tile offsets
    const int wm = wave & 1, wn = wave >> 1;
    int tm, tn;
    tile_of(blockIdx.x, tiles_m, tiles_n, tm, tn);
    int nks = (int)(K / 32);
    // a triangular factor (ktri 1: Y(j, l) = 0 for l > j; 2: X(i, l) = 0
    // for l > i): the tile's products past its diagonal 256-column k tile
    // are zero, so its k loop ends there (8 stages of 32 per k tile)
    if (ktri == 1)
        nks = min(nks, 8 * (tn + 1));
    else if (ktri == 2)
        nks = min(nks, 8 * (tm + 1));
    // strips are xstrip / ystrip bytes apart (the packed panels may be wider
    // than the K window this product uses)
    const char *xs = X + (int64_t)tm * xstrip;
    const char *ys = Y + (int64_t)tn * ystrip;

    auto issue = [&](int st) {
        const int64_t base = (int64_t)(st >> 1) * kChunk + (st & 1) * 2048;
        char *lx = smem[st % kNbuf][0];
        char *ly = smem[st % kNbuf][1];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int seg = wave * 4 + t;
            const int64_t off = base + (seg >> 1) * 4096 + (seg & 1) * 1024 + lane * 16;
            gl_lds16(xs + off, lx + seg * 1024);
            gl_lds16(ys + off, ly + seg * 1024);
        }
    };
    auto frags = [&](half8(&xd)[4], half8(&yd)[4], int st, int kk) {
        const char *lx = smem[st % kNbuf][0];
        const char *ly = smem[st % kNbuf][1];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            xd[q] = *reinterpret_cast<const half8 *>(lx + ((wm * 4 + q) * 2 + kk) * 1024 + lane * 16);
            yd[q] = *reinterpret_cast<const half8 *>(ly + ((wn * 4 + q) * 2 + kk) * 1024 + lane * 16);
        }
    };
    // copy i (0..7) of stage st: X (i even) or Y (i odd) segment wave*4 + i/2
    auto issue_one = [&](int st, int i) {
        const int64_t base = (int64_t)(st >> 1) * kChunk + (st & 1) * 2048;
        const int seg = wave * 4 + (i >> 1);
        const int64_t off = base + (seg >> 1) * 4096 + (seg & 1) * 1024 + lane * 16;
        if (i & 1)
            gl_lds16(ys + off, smem[st % kNbuf][1] + seg * 1024);
        else
            gl_lds16(xs + off, smem[st % kNbuf][0] + seg * 1024);
    };
    // fragment read i (0..7): X row i (i < 4) or Y row i - 4
    auto frag_one = [&](half8(&xd)[4], half8(&yd)[4], int st, int kk, int i) {
        if (i < 4)
            xd[i] = *reinterpret_cast<const half8 *>(smem[st % kNbuf][0] + ((wm * 4 + i) * 2 + kk) * 1024 + lane * 16);
        else
            yd[i - 4] =
                *reinterpret_cast<const half8 *>(smem[st % kNbuf][1] + ((wn * 4 + i - 4) * 2 + kk) * 1024 + lane * 16);
    };

    floatx16 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = floatx16{};

    for (int st = 0; st < kNbuf - 1 && st < nks; ++st) issue(st);
    wait_stages(min(kNbuf - 2, nks - 1));
    stage_barrier();
    half8 xa[4], ya[4], xb[4], yb[4];
    frags(xa, ya, 0, 0);
    // The LDS-DMA loads make the compiler wait for all LDS reads before an
    // MFMA that uses any of them, so fragments are read one k-step ahead, in
    // between the current MFMAs: they have landed when needed.  One stage: the first k-step's 16 MFMAs with the second k-step's
    // fragment reads between them; then the wait for the next stage and the
    // stage barrier; then the second k-step's 16 MFMAs with the refill of the
    // freed ring slot (8 LDS-DMA loads) and the next stage's first fragment
    // reads between them, so that neither the load issue nor the LDS reads
    // leave the matrix pipeline idle.  kSteady: a later stage exists and the
    // ring slot is refilled (every stage but the last kNbuf - 1).
    auto stage = [&](int st, auto steady_tag) {
        constexpr bool kSteady = decltype(steady_tag)::value;
        frags(xb, yb, st, 1);
#pragma unroll
        for (int mi = 0; mi < 4; ++mi)
#pragma unroll
            for (int ni = 0; ni < 4; ++ni)
                acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ya[ni], xa[mi], acc[mi][ni], 0, 0, 0);
#pragma unroll
        for (int g = 0; g < 8; ++g) {
            __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);  // 2 MFMA
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // 1 DS read
        }
        __builtin_amdgcn_sched_barrier(0);
        const bool more = kSteady || st + 1 < nks;
        if (kSteady) {
            wait_stages(kNbuf - 3);
            stage_barrier();
        } else if (more) {
            wait_stages(min(kNbuf - 3, nks - st - 2));
            stage_barrier();
        }
        __builtin_amdgcn_sched_barrier(0);
        if (kSteady) {
            // 8 groups of 2 MFMAs, one ring-slot copy and one fragment read,
            // in this order (fenced: left to itself the scheduler issues the
            // 8 copies back to back, each behind its M0 write, while the
            // matrix pipe idles)
            const int sn = st + kNbuf - 1;
#pragma unroll
            for (int g = 0; g < 8; ++g) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int mi = (2 * g + h) >> 2, ni = (2 * g + h) & 3;
                    acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_f16(yb[ni], xb[mi], acc[mi][ni], 0, 0, 0);
                }
                issue_one(sn, g);
                frag_one(xa, ya, st + 1, 0, g);
                __builtin_amdgcn_sched_barrier(0);
            }
            return;
        }
        if (more) {
            if (st + kNbuf - 1 < nks) issue(st + kNbuf - 1);
            frags(xa, ya, st + 1, 0);
        }
#pragma unroll
        for (int mi = 0; mi < 4; ++mi)
#pragma unroll
            for (int ni = 0; ni < 4; ++ni)
                acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_f16(yb[ni], xb[mi], acc[mi][ni], 0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
    };
    int st = 0;
    for (; st + kNbuf - 1 < nks; ++st) stage(st, std::true_type{});
    for (; st < nks; ++st) stage(st, std::false_type{});
    // per-strip power-of-two panel scales: the products carry 2^(ex + ey)
    const float a = xsexp ? ldexpf(alpha, -(xsexp[tm] + ysexp[tn])) : alpha;
    static_assert(sizeof(smem) >= 4 * 32768, "the LDS epilogue needs 32 KB per wave");
    if (!kBetaZero && c_lds_ok(M, N, C, ldc, tm, tn)) {
        stage_barrier();
        epilogue_lds<2>(acc, a, beta, C, ldc, (int64_t)tm * OMX_STRIP + wm * 128, (int64_t)tn * OMX_STRIP + wn * 128,
                        lane, wave, &smem[0][0][0]);
        return;
    }
    float mx = 0.0f;
    epilogue<kBetaZero>(acc, M, N, a, beta, C, ldc, tm, tn, wm, wn, lane, mx);
    if constexpr (kBetaZero && kTmax) {
        // the tile's max |C| (what omx_panel_tile_maxabs finds there): lanes,
        // then the four waves through the idle ring (once every wave is past
        // its last ring read).  Without kTmax the maxima above are dead code.
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
        float *red = reinterpret_cast<float *>(&smem[0][0][0]);
        __syncthreads();
        if (lane == 0) red[wave] = mx;
        __syncthreads();
        if (tid == 0)
            tmax[tmax_tr ? (int64_t)tn * tiles_m + tm : (int64_t)tm * tiles_n + tn] =
                fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    }
}

// ---------------------------------------------------------------------------
// FP8: LDS stages of 64 columns (16 KB per operand): stage s is fragment
// blocks rr*2 + (s&1) (2 KB each, two 1 KB halves) of packed chunk s/2.  One
// k-step of 16 scaled MFMAs per stage, split 8 + 8 around the barrier.
// ---------------------------------------------------------------------------
template <bool kBetaZero, int kNbuf>
__global__ void __launch_bounds__(kThreads, 1)
k_gemm_f8(int64_t M, int64_t N, int64_t K, float alpha, float beta, const char *__restrict__ X,
          const int32_t *__restrict__ xexp, const char *__restrict__ Y, const int32_t *__restrict__ yexp,
          float *__restrict__ C, int64_t ldc, int tiles_m, int tiles_n, int64_t xstrip, int64_t ystrip,
          int64_t xet, int64_t yet) {
    __shared__ __attribute__((aligned(1024))) char smem[kNbuf][2][16384];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: scalar tile offsets
    const int wm = wave & 1, wn = wave >> 1;
    int tm, tn;
    tile_of(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const int nks = (int)(K / 64);
    // strips xstrip / ystrip bytes apart; xet / yet exponents per strip
    const char *xs = X + (int64_t)tm * xstrip;
    const char *ys = Y + (int64_t)tn * ystrip;
    const int32_t *xe = xexp + (int64_t)tm * xet;
    const int32_t *ye = yexp + (int64_t)tn * yet;

    auto issue = [&](int st) {
        const int64_t base = (int64_t)(st >> 1) * kChunk + (st & 1) * 2048;
        char *lx = smem[st % kNbuf][0];
        char *ly = smem[st % kNbuf][1];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int seg = wave * 4 + t;
            const int64_t off = base + (seg >> 1) * 4096 + (seg & 1) * 1024 + lane * 16;
            gl_lds16(xs + off, lx + seg * 1024);
            gl_lds16(ys + off, ly + seg * 1024);
        }
    };
    auto frags = [&](intx8(&xd)[4], intx8(&yd)[4], int st) {
        const char *lx = smem[st % kNbuf][0];
        const char *ly = smem[st % kNbuf][1];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int fx = (wm * 4 + q) * 2048 + lane * 16;
            const int fy = (wn * 4 + q) * 2048 + lane * 16;
            const intx4 xlo = *reinterpret_cast<const intx4 *>(lx + fx);
            const intx4 xhi = *reinterpret_cast<const intx4 *>(lx + fx + 1024);
            const intx4 ylo = *reinterpret_cast<const intx4 *>(ly + fy);
            const intx4 yhi = *reinterpret_cast<const intx4 *>(ly + fy + 1024);
            xd[q] = intx8{xlo[0], xlo[1], xlo[2], xlo[3], xhi[0], xhi[1], xhi[2], xhi[3]};
            yd[q] = intx8{ylo[0], ylo[1], ylo[2], ylo[3], yhi[0], yhi[1], yhi[2], yhi[3]};
        }
    };
    // copy i (0..7) of stage st: X (i even) or Y (i odd) segment wave*4 + i/2
    auto issue_one = [&](int st, int i) {
        const int64_t base = (int64_t)(st >> 1) * kChunk + (st & 1) * 2048;
        const int seg = wave * 4 + (i >> 1);
        const int64_t off = base + (seg >> 1) * 4096 + (seg & 1) * 1024 + lane * 16;
        if (i & 1)
            gl_lds16(ys + off, smem[st % kNbuf][1] + seg * 1024);
        else
            gl_lds16(xs + off, smem[st % kNbuf][0] + seg * 1024);
    };
    // fragment i (0..7): X row i (i < 4) or Y row i - 4, both halves
    auto frag_one = [&](intx8(&xd)[4], intx8(&yd)[4], int st, int i) {
        const char *l = (i < 4 ? smem[st % kNbuf][0] + (wm * 4 + i) * 2048
                               : smem[st % kNbuf][1] + (wn * 4 + i - 4) * 2048) + lane * 16;
        const intx4 lo = *reinterpret_cast<const intx4 *>(l);
        const intx4 hi = *reinterpret_cast<const intx4 *>(l + 1024);
        const intx8 v{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
        if (i < 4)
            xd[i] = v;
        else
            yd[i - 4] = v;
    };

    floatx16 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = floatx16{};

    for (int st = 0; st < kNbuf - 1 && st < nks; ++st) issue(st);
    wait_stages(min(kNbuf - 2, nks - 1));
    stage_barrier();
    intx8 xa[4], ya[4], xb[4], yb[4];
    frags(xa, ya, 0);
    // E8M0 scales: value = fp8 * 2^(s - 127) with s = 127 - e.  One stage:
    // 8 MFMAs; the wait for the next stage and the stage barrier; 8 MFMAs
    // with the refill of the freed ring slot and the next stage's fragment
    // reads between them (kSteady: as for FP16).
    auto body = [&](intx8(&xc)[4], intx8(&yc)[4], intx8(&xn)[4], intx8(&yn)[4], int cs, int sx, int sy,
                    auto steady_tag) {
        constexpr bool kSteady = decltype(steady_tag)::value;
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int ni = 0; ni < 4; ++ni)
                acc[mi][ni] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(yc[ni], xc[mi], acc[mi][ni], 0, 0, 0,
                                                                              sy, 0, sx);
        __builtin_amdgcn_sched_barrier(0);
        const bool more = kSteady || cs + 1 < nks;
        if (kSteady) {
            wait_stages(kNbuf - 3);
            stage_barrier();
        } else if (more) {
            wait_stages(min(kNbuf - 3, nks - cs - 2));
            stage_barrier();
        }
        __builtin_amdgcn_sched_barrier(0);
        if (kSteady) {
            // 8 groups of one MFMA, one ring-slot copy and one fragment (two
            // reads), in this order (fenced, as in the FP16 kernel)
            const int sn = cs + kNbuf - 1;
#pragma unroll
            for (int g = 0; g < 8; ++g) {
                const int mi = 2 + (g >> 2), ni = g & 3;
                acc[mi][ni] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(yc[ni], xc[mi], acc[mi][ni], 0, 0, 0,
                                                                              sy, 0, sx);
                issue_one(sn, g);
                frag_one(xn, yn, cs + 1, g);
                __builtin_amdgcn_sched_barrier(0);
            }
            return;
        }
        if (more) {
            if (cs + kNbuf - 1 < nks) issue(cs + kNbuf - 1);
            frags(xn, yn, cs + 1);
        }
#pragma unroll
        for (int mi = 2; mi < 4; ++mi)
#pragma unroll
            for (int ni = 0; ni < 4; ++ni)
                acc[mi][ni] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(yc[ni], xc[mi], acc[mi][ni], 0, 0, 0,
                                                                              sy, 0, sx);
        __builtin_amdgcn_sched_barrier(0);
    };
    // nks is a multiple of 4 (k % 256 == 0): stages come in pairs, both in
    // one 256-column scale tile; the next pair's scale exponents are loaded
    // at the top of a pair, so no stage waits on a scalar load
    int st = 0;
    int sx = 127 - xe[0], sy = 127 - ye[0];
    for (; st + kNbuf < nks; st += 2) {
        const int nt = ((st + 2) * 64) / OMX_SCALE_TILE;
        const int sxn = 127 - xe[nt], syn = 127 - ye[nt];
        body(xa, ya, xb, yb, st, sx, sy, std::true_type{});
        body(xb, yb, xa, ya, st + 1, sx, sy, std::true_type{});
        sx = sxn;
        sy = syn;
    }
    for (; st < nks; st += 2) {
        const int kt = (st * 64) / OMX_SCALE_TILE;
        sx = 127 - xe[kt];
        sy = 127 - ye[kt];
        body(xa, ya, xb, yb, st, sx, sy, std::false_type{});
        body(xb, yb, xa, ya, st + 1, sx, sy, std::false_type{});
    }
    // beta != 0 interior tiles: the C update through the idle LDS ring (same
    // arithmetic as the register epilogue, bit-identical results; in the LU
    // +1 %, on its own +0.5 % at k = 10240 and +5 % at k = 5120).
    if (!kBetaZero && c_lds_ok(M, N, C, ldc, tm, tn)) {
        stage_barrier();
        static_assert(sizeof(smem) >= 4 * 32768, "the LDS epilogue needs 32 KB per wave");
        epilogue_lds<4>(acc, alpha, beta, C, ldc, (int64_t)tm * OMX_STRIP + wm * 128,
                     (int64_t)tn * OMX_STRIP + wn * 128, lane, wave, &smem[0][0][0]);
        return;
    }
    float mx = 0.0f;  // (unused here)
    epilogue<kBetaZero>(acc, M, N, alpha, beta, C, ldc, tm, tn, wm, wn, lane, mx);
}

// ---------------------------------------------------------------------------
// Two workgroups per CU.  Workgroup tile 256 x 128: an X strip against one
// half (128 rows) of a Y strip.  Four waves in a 2 x 2 arrangement, each
// owning 128 x 64 of C as 4 x 2 accumulators of a 32 x 32 MFMA (128
// registers), so that a wave fits in 256 registers and two workgroups share
// a CU: while one waits on memory, passes a barrier or runs its C epilogue,
// the other keeps the matrix pipes busy.
//
// LDS: a ring of kSlots slots of 12 KB (72 KB per workgroup).  Slot s holds,
// for each of the 8 X and 4 Y fragment rows (32 matrix rows each), the 1 KB
// block (s & 3) of packed chunk s >> 2: in FP16 the 32 x 16 fragment of
// columns 16 s .. 16 s + 15, in FP8 one 1 KB half of a 32 x 64 fragment.  A
// k-step (one MFMA depth: 16 FP16 or 64 FP8 columns) takes kP slots (1 or
// 2); the ring holds kR = kSlots / kP k-steps.
//
// Step st: wait until this wave's copies of step st + 1 have landed and its
// fragment reads of step st are in registers; barrier (so every wave's
// copies of step st + 1 are in LDS and no wave still reads step st's
// slots); refill step st's slots with step st + kR; read step st + 1's
// fragments; run step st's 8 MFMAs.  A copy has kR - 1 steps to land.
// ---------------------------------------------------------------------------
constexpr int kSlots = 6;
constexpr int kSlotBytes = 12288;

// s_waitcnt vmcnt(v) lgkmcnt(0) (expcnt not waited), v < 64
template <int v>
__device__ __forceinline__ void wait_vm_lgkm0() {
    static_assert(v >= 0 && v < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((v & 15) | ((v >> 4) << 14) | (7 << 4));
}

// wait until at most c k-steps of copies (kLoads per wave each) are in
// flight and every LDS read has returned
template <int kLoads>
__device__ __forceinline__ void wait_steps(int c) {
    switch (c) {
        case 0: wait_vm_lgkm0<0>(); break;
        case 1: wait_vm_lgkm0<kLoads>(); break;
        case 2: wait_vm_lgkm0<2 * kLoads>(); break;
        case 3: wait_vm_lgkm0<3 * kLoads>(); break;
        case 4: wait_vm_lgkm0<4 * kLoads>(); break;
        default: wait_vm_lgkm0<5 * kLoads>(); break;
    }
}

// C update of one wave's 128 x 64 block: acc[mi][ni][r] is
//   C(i0 + 32 mi + lane%32, j0 + 32 ni + 8 (r/4) + 4 (lane/32) + r%4).
// Interior tiles: buffer instructions, one resource per 32-column block; the
// old values move in units of two row blocks of one column block (32
// registers), the next unit's loads in flight while one is combined and
// stored.
template <bool kBetaZero>
__device__ __forceinline__ void epilogue_half(floatx16 (&acc)[4][2], int64_t M, int64_t N, float alpha, float beta,
                                              float *__restrict__ C, int64_t ldc, int tm, int tn, int wm, int wn,
                                              int lane) {
    const int64_t i0 = (int64_t)tm * OMX_STRIP + wm * 128;
    const int64_t j0 = (int64_t)tn * 128 + wn * 64;
    if ((int64_t)(tm + 1) * OMX_STRIP <= M && (int64_t)(tn + 1) * 128 <= N && ldc * 32 * 4 < ((int64_t)1 << 31)) {
        int voff[16];
#pragma unroll
        for (int r = 0; r < 16; ++r)
            voff[r] = (int)(((8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)) * ldc + (lane & 31)) * 4);
        if (kBetaZero) {
#pragma unroll
            for (int ni = 0; ni < 2; ++ni) {
                const __amdgpu_buffer_rsrc_t rs = colblock_rsrc(C + (j0 + 32 * ni) * ldc + i0, ldc);
#pragma unroll
                for (int mi = 0; mi < 4; ++mi)
#pragma unroll
                    for (int r = 0; r < 16; ++r)
                        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, alpha * acc[mi][ni][r]),
                                                              rs, voff[r] + mi * 128, 0, kCPolicy);
            }
            return;
        }
        float old[2][2][16];
        auto fetch = [&](float(&o)[2][16], int u) {
            const int ni = u >> 1, m0 = 2 * (u & 1);
            const __amdgpu_buffer_rsrc_t rs = colblock_rsrc(C + (j0 + 32 * ni) * ldc + i0, ldc);
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    o[h][r] = __builtin_bit_cast(
                        float, __builtin_amdgcn_raw_buffer_load_b32(rs, voff[r] + (m0 + h) * 128, 0, kCPolicy));
        };
        fetch(old[0], 0);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (u + 1 < 4) fetch(old[(u + 1) & 1], u + 1);
            const int ni = u >> 1, m0 = 2 * (u & 1);
            const __amdgpu_buffer_rsrc_t rs = colblock_rsrc(C + (j0 + 32 * ni) * ldc + i0, ldc);
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    __builtin_amdgcn_raw_buffer_store_b32(
                        __builtin_bit_cast(uint32_t, beta * old[u & 1][h][r] + alpha * acc[m0 + h][ni][r]), rs,
                        voff[r] + (m0 + h) * 128, 0, kCPolicy);
        }
        return;
    }
    // edge tiles (and ldc too large for 32-bit column offsets): checked
    const int64_t ib = i0 + (lane & 31);
    const int64_t jb = j0 + 4 * (lane >> 5);
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
        const int64_t i = ib + mi * 32;
        if (i >= M) continue;
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) {
            float *cp = C + i;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t j = jb + ni * 32 + 8 * (r >> 2) + (r & 3);
                if (j < N) {
                    const float v = acc[mi][ni][r];
                    cp[j * ldc] = kBetaZero ? alpha * v : beta * cp[j * ldc] + alpha * v;
                }
            }
        }
    }
}

template <bool kMore_, bool kIssue_, int kWait_>
struct StepCfg {
    static constexpr bool kMore = kMore_;
    static constexpr bool kIssue = kIssue_;
    static constexpr int kWait = kWait_;
};

// The last kT steps (kT even) from step s0, alternating the two fragment
// sets; step t's flags from its position.
template <int kT, int kR, int t = 0, class F, class XA, class YA>
__device__ __forceinline__ void tail_steps(F &step, XA &xa, YA &ya, XA &xb, YA &yb, int s0) {
    if constexpr (t < kT) {
        constexpr int kWait = (kR - 2 < kT - t - 2) ? kR - 2 : kT - t - 2;
        typedef StepCfg<(t < kT - 1), (t + kR < kT), (kWait < 0 ? 0 : kWait)> Cfg;
        if constexpr (t % 2 == 0)
            step(xa, ya, xb, yb, s0 + t, Cfg{});
        else
            step(xb, yb, xa, ya, s0 + t, Cfg{});
        tail_steps<kT, kR, t + 1>(step, xa, ya, xb, yb, s0);
    }
}

// Step depth and fragment buffering of the half-tile kernels: kP slots per
// k-step (FP16: 1 or 2 16-column MFMA depths per step; FP8: always 2, the
// two halves of one 64-column operand); kDB: the next step's fragments in
// their own registers, read between the current step's MFMAs (double
// buffering), or (0) each fragment row read into the registers of the one
// it replaces once that one's last MFMA has issued.
#ifndef OMX_HALF_F16_P
#define OMX_HALF_F16_P 1
#endif
#ifndef OMX_HALF_F16_DB
#define OMX_HALF_F16_DB 1
#endif
#ifndef OMX_HALF_F8_DB
#define OMX_HALF_F8_DB 1
#endif

template <bool kFp8, bool kBetaZero, int kP, bool kDB>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_gemm_half(int64_t M, int64_t N, int64_t K, float alpha, float beta, const char *__restrict__ X,
            const int32_t *__restrict__ xexp, const char *__restrict__ Y, const int32_t *__restrict__ yexp,
            float *__restrict__ C, int64_t ldc, int tiles_m, int tiles_n, int64_t xstrip, int64_t ystrip,
            int64_t xet, int64_t yet) {
    static_assert(!kFp8 || kP == 2, "an FP8 operand spans two slots");
    constexpr int kR = kSlots / kP;                    // k-steps in the ring
    constexpr int kCols = (kFp8 ? 32 : 16) * kP;       // columns per k-step
    constexpr int kLoads = 3 * kP;                     // copies per wave per k-step
    typedef typename std::conditional<kP == 1, intx4, intx8>::type Frag;
    __shared__ __attribute__((aligned(1024))) char smem[kSlots * kSlotBytes];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: scalar tile offsets
    const int wm = wave & 1, wn = wave >> 1;
    int tm, tn;
    tile_of(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const int nsteps = (int)(K / kCols);
    const int ys_strip = tn >> 1;
    // this wave's three 1 KB blocks of every slot: block b = 3 wave + t is X
    // fragment row b (b < 8) or Y fragment row 4 (tn & 1) + b - 8
    const char *src[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const int b = 3 * wave + t;
        src[t] = (b < 8 ? X + (int64_t)tm * xstrip + b * 4096
                        : Y + (int64_t)ys_strip * ystrip + (4 * (tn & 1) + b - 8) * 4096) + lane * 16;
    }
    const int32_t *xe = kFp8 ? xexp + (int64_t)tm * xet : nullptr;
    const int32_t *ye = kFp8 ? yexp + (int64_t)ys_strip * yet : nullptr;

    // k-step `step` is slots kP step .. kP step + kP - 1
    auto issue = [&](int step) {
#pragma unroll
        for (int p = 0; p < kP; ++p) {
            const int s = kP * step + p;
            const int64_t off = (int64_t)(s >> 2) * kChunk + (s & 3) * 1024;
            char *l = smem + (s % kSlots) * kSlotBytes;
#pragma unroll
            for (int t = 0; t < 3; ++t) gl_lds16(src[t] + off, l + (3 * wave + t) * 1024);
        }
    };
    // copy i (0 .. 3 kP - 1) of k-step `step`: slot kP step + i / 3, block
    // 3 wave + i % 3
    auto issue_one = [&](int step, int i) {
        const int s = kP * step + i / 3, t = i % 3;
        const int64_t off = (int64_t)(s >> 2) * kChunk + (s & 3) * 1024;
        gl_lds16(src[t] + off, smem + (s % kSlots) * kSlotBytes + (3 * wave + t) * 1024);
    };
    // A lane's operand of one k-step for fragment row f (0-7 X, 8-11 Y): 16
    // bytes from each of the step's slots.  FP16: one 16-column MFMA operand
    // per slot; FP8: the 32 bytes of one 64-column MFMA operand.
    auto frag = [&](int step, int f) -> Frag {
        const char *l0 = smem + ((kP * step) % kSlots) * kSlotBytes + f * 1024 + lane * 16;
        const intx4 lo = *reinterpret_cast<const intx4 *>(l0);
        if constexpr (kP == 1) {
            return lo;
        } else {
            const intx4 hi = *reinterpret_cast<const intx4 *>(l0 + kSlotBytes);
            return intx8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
        }
    };
    auto mfma = [&](floatx16 &c, const Frag &y, const Frag &x, int sy, int sx) {
        if constexpr (kFp8) {
            c = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(y, x, c, 0, 0, 0, sy, 0, sx);
        } else if constexpr (kP == 1) {
            c = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(half8, y), __builtin_bit_cast(half8, x), c,
                                                       0, 0, 0);
        } else {
            const intx4 ylo{y[0], y[1], y[2], y[3]}, yhi{y[4], y[5], y[6], y[7]};
            const intx4 xlo{x[0], x[1], x[2], x[3]}, xhi{x[4], x[5], x[6], x[7]};
            c = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(half8, ylo), __builtin_bit_cast(half8, xlo),
                                                       c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(half8, yhi), __builtin_bit_cast(half8, xhi),
                                                       c, 0, 0, 0);
        }
    };

    floatx16 acc[4][2];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = floatx16{};

    for (int st = 0; st < kR && st < nsteps; ++st) issue(st);
    wait_steps<kLoads>(min(kR - 1, nsteps - 1));
    stage_barrier();
    Frag xa[4], ya[2], xb[4], yb[2];
#pragma unroll
    for (int q = 0; q < 4; ++q) xa[q] = frag(0, wm * 4 + q);
#pragma unroll
    for (int q = 0; q < 2; ++q) ya[q] = frag(0, 8 + wn * 2 + q);

    // One k-step.  Cfg::kMore: a later step exists (its fragments are read);
    // Cfg::kIssue: step st + kR exists and is copied into step st's slots;
    // Cfg::kWait: k-steps of copies that may stay in flight at the wait.
    auto step = [&](Frag(&xc)[4], Frag(&yc)[2], Frag(&xn)[4], Frag(&yn)[2], int st, auto cfg) {
        typedef decltype(cfg) Cfg;
        constexpr bool more = Cfg::kMore;
        if (more) {
            wait_steps<kLoads>(Cfg::kWait);
            stage_barrier();
            if (!kDB && Cfg::kIssue) issue(st + kR);
        }
        int sx = 0, sy = 0;
        if constexpr (kFp8) {
            const int kt = (st * 64) / OMX_SCALE_TILE;
            sx = 127 - xe[kt];
            sy = 127 - ye[kt];
        }
        if constexpr (kDB) {
            // the step's 8 (mi, ni) products in groups, each followed by one
            // of the refill copies and one of the next step's fragments
            // (fenced: left to itself the scheduler issues the copies back to
            // back, each behind its M0 write, while the matrix pipe idles)
#pragma unroll
            for (int g = 0; g < 8; ++g) {
                mfma(acc[g >> 1][g & 1], yc[g & 1], xc[g >> 1], sy, sx);
                if (Cfg::kIssue && g < 3 * kP) issue_one(st + kR, g);
                if (more && g < 6) {
                    if (g < 4)
                        xn[g] = frag(st + 1, wm * 4 + g);
                    else
                        yn[g - 4] = frag(st + 1, 8 + wn * 2 + g - 4);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        } else {
            // the next step's X fragment of row mi is read once the current
            // one's last MFMA has issued, its Y fragments after the last
            // MFMA (fragments in 2/3 of the double-buffered registers)
#pragma unroll
            for (int mi = 0; mi < 4; ++mi) {
#pragma unroll
                for (int ni = 0; ni < 2; ++ni) mfma(acc[mi][ni], yc[ni], xc[mi], sy, sx);
                if (more) xn[mi] = frag(st + 1, wm * 4 + mi);
            }
            if (more) {
                yn[0] = frag(st + 1, 8 + wn * 2);
                yn[1] = frag(st + 1, 8 + wn * 2 + 1);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    };
    // nsteps is even.  Steady state: pairs of steps with every flag set;
    // then the last kT steps with their flags fixed at compile time (tail
    // step t: a later step if t < kT - 1, a refill if t + kR < kT), so
    // neither part branches.  FP8 products have at least 4 steps (K % 256
    // == 0); shorter FP16 products (K = 64: 4 steps of 16 or 2 of 32) take
    // a static tail of their own length.
    constexpr int kT = kR + 1 + ((kR + 1) & 1);  // even, > kR
    if (kFp8 || nsteps >= kT) {
        int st = 0;
        for (; st + kT < nsteps; st += 2) {
            step(xa, ya, xb, yb, st, StepCfg<true, true, kR - 2>{});
            step(xb, yb, xa, ya, st + 1, StepCfg<true, true, kR - 2>{});
        }
        tail_steps<kT, kR>(step, xa, ya, xb, yb, nsteps - kT);
    } else if constexpr (!kFp8) {
        constexpr int kShort = 64 / kCols;  // K = 64
        tail_steps<kShort, kR>(step, xa, ya, xb, yb, 0);
    }
    float a = alpha;
    if (!kFp8 && xexp) a = ldexpf(alpha, -(xexp[tm] + yexp[ys_strip]));
    epilogue_half<kBetaZero>(acc, M, N, a, beta, C, ldc, tm, tn, wm, wn, lane);
}

constexpr int kNbuf = OMX_GEMM_NBUF;

int gemm_args(int64_t m, int64_t n, int64_t k, int64_t kstage, const void *X, int64_t xr0,
              const void *Y, int64_t yr0, const float *C, int64_t ldc) {
    if (m < 0 || n < 0 || k <= 0 || k % kstage || !X || !Y || !C || ldc < m) return OMX_ERR_ARG;
    if (xr0 < 0 || yr0 < 0 || xr0 % OMX_STRIP || yr0 % OMX_STRIP) return OMX_ERR_ARG;
    const int64_t tiles = omx_cdiv(m, OMX_STRIP) * omx_cdiv(n, OMX_STRIP);
    if (tiles >= ((int64_t)1 << 31)) return OMX_ERR_ARG;
    return OMX_OK;
}

// The trailing-update kernel family: 256 x 256 tiles one workgroup per CU
// (default) or 256 x 128 tiles two per CU (OMX_GEMM_TILE=half).  The second
// is faster on a GPU that runs nothing else (FP8 2.4-2.5 against 2.35-2.4
// PFLOP/s at k = 10240), the first inside the LU, where the critical
// stream's kernels run beside the trailing update (FP8 1.94 against 1.60,
// FP16 1.02 against 0.87 PFLOP/s).
bool use_half_tiles() {
    const char *e = getenv("OMX_GEMM_TILE");
    return e && !strcmp(e, "half");
}

// the K window [k0, k0 + k) of a packed panel of ldk columns
bool window_ok(int64_t k, int64_t ldk, int64_t k0, int64_t align) {
    return ldk > 0 && ldk % align == 0 && k0 >= 0 && k0 % align == 0 && k0 + k <= ldk;
}

// The FP16 products; tri and tile_max as in omx_gemm_f16_packed_tri (0 and
// NULL: full K, no tile maxima).
int gemm_f16_launch(int64_t m, int64_t n, int64_t k, float alpha, const uint16_t *X, const int32_t *xsexp,
                    int64_t x_row0, int64_t x_ldk, int64_t x_k0, const uint16_t *Y, const int32_t *ysexp,
                    int64_t y_row0, int64_t y_ldk, int64_t y_k0, float beta, float *C, int64_t ldc, int tri,
                    float *tile_max, void *stream) {
    int rc = gemm_args(m, n, k, OMX_KSTAGE16, X, x_row0, Y, y_row0, C, ldc);
    if (rc) return rc;
    if ((xsexp == nullptr) != (ysexp == nullptr)) return OMX_ERR_ARG;
    if (!window_ok(k, x_ldk, x_k0, OMX_KSTAGE16) || !window_ok(k, y_ldk, y_k0, OMX_KSTAGE16)) return OMX_ERR_ARG;
    if (m == 0 || n == 0) return OMX_OK;
    const int tiles_m = (int)omx_cdiv(m, OMX_STRIP), tiles_n = (int)omx_cdiv(n, OMX_STRIP);
    const int64_t xstrip = (x_ldk / OMX_KSTAGE16) * kChunk, ystrip = (y_ldk / OMX_KSTAGE16) * kChunk;
    const char *xp = reinterpret_cast<const char *>(X) + (x_row0 / OMX_STRIP) * xstrip + (x_k0 / OMX_KSTAGE16) * kChunk;
    const char *yp = reinterpret_cast<const char *>(Y) + (y_row0 / OMX_STRIP) * ystrip + (y_k0 / OMX_KSTAGE16) * kChunk;
    const int32_t *xe = xsexp ? xsexp + x_row0 / OMX_STRIP : nullptr;
    const int32_t *ye = ysexp ? ysexp + y_row0 / OMX_STRIP : nullptr;
    if (use_half_tiles() && !tile_max) {
        // (full K also for a triangular factor: its skipped products are zero)
        const int tiles_h = (int)omx_cdiv(n, 128);
        if ((int64_t)tiles_m * tiles_h >= ((int64_t)1 << 31)) return OMX_ERR_ARG;
        dim3 grid((unsigned)(tiles_m * tiles_h));
        if (beta == 0.0f)
            hipLaunchKernelGGL((k_gemm_half<false, true, OMX_HALF_F16_P, OMX_HALF_F16_DB>), grid, dim3(kThreads), 0, omx_stream(stream), m, n, k,
                               alpha, beta, xp, xe, yp, ye, C, ldc, tiles_m, tiles_h, xstrip, ystrip, (int64_t)0,
                               (int64_t)0);
        else
            hipLaunchKernelGGL((k_gemm_half<false, false, OMX_HALF_F16_P, OMX_HALF_F16_DB>), grid, dim3(kThreads), 0, omx_stream(stream), m, n, k,
                               alpha, beta, xp, xe, yp, ye, C, ldc, tiles_m, tiles_h, xstrip, ystrip, (int64_t)0,
                               (int64_t)0);
        OMX_LAUNCH_CHECK();
        return OMX_OK;
    }
    dim3 grid((unsigned)(tiles_m * tiles_n));
    if (beta == 0.0f && tile_max)
        hipLaunchKernelGGL((k_gemm_f16<true, kNbuf, true>), grid, dim3(kThreads), 0, omx_stream(stream), m, n, k,
                           alpha, beta, xp, yp, C, ldc, tiles_m, tiles_n, xe, ye, xstrip, ystrip, tri, tile_max,
                           tri == 2 ? 1 : 0);
    else if (beta == 0.0f)
        hipLaunchKernelGGL((k_gemm_f16<true, kNbuf>), grid, dim3(kThreads), 0, omx_stream(stream), m, n, k, alpha,
                           beta, xp, yp, C, ldc, tiles_m, tiles_n, xe, ye, xstrip, ystrip, tri, nullptr, 0);
    else
        hipLaunchKernelGGL((k_gemm_f16<false, kNbuf>), grid, dim3(kThreads), 0, omx_stream(stream), m, n, k, alpha,
                           beta, xp, yp, C, ldc, tiles_m, tiles_n, xe, ye, xstrip, ystrip, tri, nullptr, 0);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

}  // namespace

extern "C" {

int omx_gemm_f16_packed_window(int64_t m, int64_t n, int64_t k, float alpha, const uint16_t *X,
                               const int32_t *xsexp, int64_t x_row0, int64_t x_ldk, int64_t x_k0,
                               const uint16_t *Y, const int32_t *ysexp, int64_t y_row0, int64_t y_ldk,
                               int64_t y_k0, float beta, float *C, int64_t ldc, void *stream) {
    return gemm_f16_launch(m, n, k, alpha, X, xsexp, x_row0, x_ldk, x_k0, Y, ysexp, y_row0, y_ldk, y_k0, beta, C, ldc,
                           0, nullptr, stream);
}

int omx_gemm_f16_packed_tri(int64_t m, int64_t n, int64_t k, float alpha, const uint16_t *X, const uint16_t *Y,
                            int tri, float beta, float *C, int64_t ldc, float *tile_max, void *stream) {
    if (tri < 0 || tri > 2 || (tile_max && beta != 0.0f)) return OMX_ERR_ARG;
    return gemm_f16_launch(m, n, k, alpha, X, nullptr, 0, k, 0, Y, nullptr, 0, k, 0, beta, C, ldc, tri, tile_max,
                           stream);
}

int omx_gemm_f16_packed_scaled(int64_t m, int64_t n, int64_t k, float alpha, const uint16_t *X,
                               const int32_t *xsexp, int64_t x_row0, const uint16_t *Y, const int32_t *ysexp,
                               int64_t y_row0, float beta, float *C, int64_t ldc, void *stream) {
    return omx_gemm_f16_packed_window(m, n, k, alpha, X, xsexp, x_row0, k, 0, Y, ysexp, y_row0, k, 0, beta, C, ldc,
                                      stream);
}

int omx_gemm_f16_packed(int64_t m, int64_t n, int64_t k, float alpha, const uint16_t *X,
                        int64_t x_row0, const uint16_t *Y, int64_t y_row0, float beta, float *C,
                        int64_t ldc, void *stream) {
    return omx_gemm_f16_packed_window(m, n, k, alpha, X, nullptr, x_row0, k, 0, Y, nullptr, y_row0, k, 0, beta, C,
                                      ldc, stream);
}

int omx_gemm_f8_packed_window(int64_t m, int64_t n, int64_t k, float alpha, const uint8_t *X, const int32_t *xexp,
                              int64_t x_row0, int64_t x_ldk, int64_t x_k0, const uint8_t *Y, const int32_t *yexp,
                              int64_t y_row0, int64_t y_ldk, int64_t y_k0, float beta, float *C, int64_t ldc,
                              void *stream) {
    int rc = gemm_args(m, n, k, OMX_SCALE_TILE, X, x_row0, Y, y_row0, C, ldc);
    if (rc) return rc;
    if (!xexp || !yexp) return OMX_ERR_ARG;
    if (!window_ok(k, x_ldk, x_k0, OMX_SCALE_TILE) || !window_ok(k, y_ldk, y_k0, OMX_SCALE_TILE)) return OMX_ERR_ARG;
    if (m == 0 || n == 0) return OMX_OK;
    const int tiles_m = (int)omx_cdiv(m, OMX_STRIP), tiles_n = (int)omx_cdiv(n, OMX_STRIP);
    const int64_t xstrip = (x_ldk / OMX_KSTAGE8) * kChunk, ystrip = (y_ldk / OMX_KSTAGE8) * kChunk;
    const int64_t xet = x_ldk / OMX_SCALE_TILE, yet = y_ldk / OMX_SCALE_TILE;
    const char *xp = reinterpret_cast<const char *>(X) + (x_row0 / OMX_STRIP) * xstrip + (x_k0 / OMX_KSTAGE8) * kChunk;
    const char *yp = reinterpret_cast<const char *>(Y) + (y_row0 / OMX_STRIP) * ystrip + (y_k0 / OMX_KSTAGE8) * kChunk;
    const int32_t *xe = xexp + (x_row0 / OMX_STRIP) * xet + x_k0 / OMX_SCALE_TILE;
    const int32_t *ye = yexp + (y_row0 / OMX_STRIP) * yet + y_k0 / OMX_SCALE_TILE;
    if (use_half_tiles()) {
        const int tiles_h = (int)omx_cdiv(n, 128);
        if ((int64_t)tiles_m * tiles_h >= ((int64_t)1 << 31)) return OMX_ERR_ARG;
        dim3 grid((unsigned)(tiles_m * tiles_h));
        if (beta == 0.0f)
            hipLaunchKernelGGL((k_gemm_half<true, true, 2, OMX_HALF_F8_DB>), grid, dim3(kThreads), 0, omx_stream(stream), m, n, k,
                               alpha, beta, xp, xe, yp, ye, C, ldc, tiles_m, tiles_h, xstrip, ystrip, xet, yet);
        else
            hipLaunchKernelGGL((k_gemm_half<true, false, 2, OMX_HALF_F8_DB>), grid, dim3(kThreads), 0, omx_stream(stream), m, n, k,
                               alpha, beta, xp, xe, yp, ye, C, ldc, tiles_m, tiles_h, xstrip, ystrip, xet, yet);
        OMX_LAUNCH_CHECK();
        return OMX_OK;
    }
    dim3 grid((unsigned)(tiles_m * tiles_n));
    if (beta == 0.0f)
        hipLaunchKernelGGL((k_gemm_f8<true, kNbuf>), grid, dim3(kThreads), 0, omx_stream(stream), m, n, k, alpha,
                           beta, xp, xe, yp, ye, C, ldc, tiles_m, tiles_n, xstrip, ystrip, xet, yet);
    else
        hipLaunchKernelGGL((k_gemm_f8<false, kNbuf>), grid, dim3(kThreads), 0, omx_stream(stream), m, n, k, alpha,
                           beta, xp, xe, yp, ye, C, ldc, tiles_m, tiles_n, xstrip, ystrip, xet, yet);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

int omx_gemm_f8_packed(int64_t m, int64_t n, int64_t k, float alpha, const uint8_t *X,
                       const int32_t *xexp, int64_t x_row0, const uint8_t *Y, const int32_t *yexp,
                       int64_t y_row0, float beta, float *C, int64_t ldc, void *stream) {
    return omx_gemm_f8_packed_window(m, n, k, alpha, X, xexp, x_row0, k, 0, Y, yexp, y_row0, k, 0, beta, C, ldc,
                                     stream);
}

}  // extern "C"
