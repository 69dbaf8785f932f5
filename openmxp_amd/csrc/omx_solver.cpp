// omx_solver.cpp -- host driver of the solver: blocked no-pivot LU with
// look-ahead on two HIP streams, and FP64 iterative refinement (classical
// and GMRES) on the GPU.
//
// LU (panel_lu_async_wGPU_gdirect_alt2, panel_trf_gdirect.hpp:593-1192, on a
// 1x1 grid).  Panel i is the factored diagonal block A_ii with its column
// panel A[i+1:, i] and row panel A[i, i+1:], packed to FP16 or FP8 as L_i
// and R_i.  Every block (I, J) with I, J > i receives -L_i[I] R_i[J]^T, in
// the order i = 0, 1, 2, ...; the reference applies one panel per pass over
// the trailing matrix.  Here the trailing matrix receives the panels in
// groups of G (2, or 4 with OMX_SOLVER_GROUP4): the FP32 trailing matrix is
// read and written once per group with one product of depth G b instead of
// G times (its traffic, not the MFMA work, bounds the update's epilogue).
// The panels of a group share one packed buffer (panel G g + j in K columns
// [j b, (j+1) b)).  Group g, with panels G g .. G g + G - 1 factored:
//   bulk stream  block row/col t -= group g for each block t of the next
//                group (the look-ahead products), then
//                A[G(g+2):, G(g+2):] -= [L_Gg .. L_Gg+G-1] [R_Gg .. R_Gg+G-1]^T
//   crit stream  for each block t of the next group, in order:
//                block row/col t -= the next group's panels before t;
//                factor panel t
// The crit stream waits for the look-ahead products of group g (queued on
// the bulk stream after the trailing update of group g-1, which covered the
// next group's blocks), the bulk stream for group g's panels; so the
// factorisation of the next group's panels overlaps the trailing update of
// the current group, and the look-ahead products run beside the
// factorisation chain rather than beside the trailing update.  Runs of a group's panels packed in the same
// precision are applied as one product; panels in scaled FP16 (per-panel
// scales) one by one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "omx_common.hpp"
#include "omx_layout.hpp"

#define RC(x)                  \
    do {                       \
        int rc_ = (x);         \
        if (rc_) return rc_;   \
    } while (0)

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

int dalloc(DevBuf &b, size_t bytes) {
    b.bytes = bytes;
    if (bytes == 0) return OMX_OK;
    return hipMalloc(&b.p, bytes) == hipSuccess ? OMX_OK : OMX_ERR_HIP;
}

void dfree(DevBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
}

// Phase timers (OMX_SOLVER_TIMERS): each timed phase records a pair of
// events from a reusable pool on the stream that runs it; after the call
// synchronises, the pairs' elapsed times are summed per item.  Items follow
// the reference's Timer enumeration (timer.hpp:57-81).
struct PhaseTimers {
    bool on = false;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    struct Span {
        int item;
        size_t e0, e1;
        double ops;
    };
    std::vector<Span> spans;
    double seconds[OMX_TIMER_ITEMS] = {}, ops[OMX_TIMER_ITEMS] = {};

    int event(size_t &idx) {
        if (used == pool.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return OMX_ERR_HIP;
            pool.push_back(e);
        }
        idx = used++;
        return OMX_OK;
    }
    // start a phase on st; returns the span index through *span (-1 when off)
    int beg(int item, hipStream_t st, long *span) {
        *span = -1;
        if (!on) return OMX_OK;
        size_t e0;
        RC(event(e0));
        if (hipEventRecord(pool[e0], st) != hipSuccess) return OMX_ERR_HIP;
        spans.push_back(Span{item, e0, e0, 0.0});
        *span = (long)spans.size() - 1;
        return OMX_OK;
    }
    int end(long span, hipStream_t st, double op) {
        if (!on || span < 0) return OMX_OK;
        size_t e1;
        RC(event(e1));
        if (hipEventRecord(pool[e1], st) != hipSuccess) return OMX_ERR_HIP;
        spans[span].e1 = e1;
        spans[span].ops = op;
        return OMX_OK;
    }
    // after the streams are synchronised
    int collect() {
        for (const Span &sp : spans) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, pool[sp.e0], pool[sp.e1]) != hipSuccess) return OMX_ERR_HIP;
            seconds[sp.item] += ms * 1e-3;
            ops[sp.item] += sp.ops;
        }
        spans.clear();
        used = 0;
        return OMX_OK;
    }
    void reset() {
        for (int i = 0; i < OMX_TIMER_ITEMS; ++i) seconds[i] = ops[i] = 0.0;
        spans.clear();
        used = 0;
    }
};

// Run f (returning an OMX code) as phase `item` on stream st with `op`
// operations.
#define TIMED(T, item, st, op, f)                  \
    do {                                           \
        long sp_;                                  \
        RC((T).beg((item), (st), &sp_));           \
        RC(f);                                     \
        RC((T).end(sp_, (st), (op)));              \
    } while (0)

// the reference's Timer item numbers (timer.hpp:57-81) of the timed phases
enum {
    kT_DIAG_LU = 5,
    kT_TRSM_L = 6,
    kT_TRSM_R = 7,
    kT_CONV_L = 8,
    kT_CONV_R = 9,
    kT_GEMM_UPDATE = 10,
    kT_IR_GEMV = 14,
    kT_IR_TRSV = 16,
    kT_TOTAL = 21,
};

}  // namespace

struct omx_solver {
    int64_t n = 0, b = 0, nb = 0;
    int alt = 1;
    int fp8 = 0;
    hipStream_t crit = nullptr, bulk = nullptr;
    std::vector<hipEvent_t> ev_panel, ev_t0, ev_t1;
    int64_t G = 2;                      // panels per trailing update (group)
    DevBuf Lp[2], Rp[2], Tp;            // FP16 group buffers (K = G b) and the -alt temporary
    DevBuf L8[2], R8[2], Le[2], Re[2];  // FP8 group buffers and their exponents
    DevBuf L8max, R8max;                // tile max-abs scratch
    int scale16 = 0;                    // FP16 Schur panels with per-strip scales
    DevBuf Lse[8], Rse[8];              // their strip exponents, per (group buffer, window)
    DevBuf Wup, Wlo, Wup_h, Wlo_h;      // -alt inverses (FP32, packed FP16)
    DevBuf Wtmp;                        // their b x b scratch
    DevBuf info;
    // learned schedule: MLP parameters, per-step features, forward scratch
    DevBuf pol, feat, mlp_h, zdev;
    int has_policy = 0;
    hipEvent_t ev_dec = nullptr;
    std::vector<hipEvent_t> ev_la;     // per block of the next group: its look-ahead products done
    float *zhost = nullptr;            // pinned: the step's 3 outputs
    std::vector<int> last_sched;       // per step: 0 FP16, 1 + headroom FP8
    std::vector<int> prec;             // the current factorisation's choices
    std::vector<float> last_feat;      // per step: 16 features (FP8 solvers)
    // refinement
    DevBuf r, w, gemv_ws, trsv_ws, nrm, V, hv, dots;
    int restart_cap = 0;
    // the reference's -skip bits 1-3 (main.cpp:68-70): leave out the
    // diagonal getrf, the panel solves / inverse products, the Schur products
    int skip = 0;
    // progress: an event at the start of the factorisation and one per group
    // when all of the group's Schur products are done
    hipEvent_t ev_start = nullptr;
    std::vector<hipEvent_t> ev_grp;
    int64_t ngrp_done = 0;
    // timings of the last factorisation
    double t_lu_ms = 0, bulk_ms = 0, bulk_flops = 0;
    int bulk_count = 0;
    double crit_flops = 0;
    PhaseTimers timers;
};

namespace {

constexpr int kDotBlocks = 512;

__global__ void k_dot_partial(int64_t n, int nv, const double *__restrict__ V, int64_t ldv,
                              const double *__restrict__ w, double *__restrict__ part) {
    // part[v * gridDim.x + block] = sum over this block's slice of V[:,v] * w
    __shared__ double sh[256];
    const int64_t chunk = omx_cdiv(n, gridDim.x);
    const int64_t i0 = (int64_t)blockIdx.x * chunk;
    const int64_t i1 = (i0 + chunk < n) ? i0 + chunk : n;
    for (int v = 0; v < nv; ++v) {
        const double *col = V + (int64_t)v * ldv;
        double s = 0.0;
        for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) s += col[i] * w[i];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(int64_t)v * gridDim.x + blockIdx.x] = sh[0];
        __syncthreads();
    }
}

__global__ void k_dot_final(int nv, int nparts, const double *__restrict__ part, double *__restrict__ out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double s = 0.0;
    for (int q = 0; q < nparts; ++q) s += part[(int64_t)v * nparts + q];
    out[v] = s;
}

// w -= V[:, 0:nv] * h  (h on the device)
__global__ void k_multi_axpy(int64_t n, int nv, const double *__restrict__ V, int64_t ldv,
                             const double *__restrict__ h, double *__restrict__ w) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double s = w[i];
        for (int v = 0; v < nv; ++v) s -= V[i + (int64_t)v * ldv] * h[v];
        w[i] = s;
    }
}

// x += V[:, 0:nv] * y
__global__ void k_multi_axpy_add(int64_t n, int nv, const double *__restrict__ V, int64_t ldv,
                                 const double *__restrict__ y, double *__restrict__ x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int v = 0; v < nv; ++v) s += V[i + (int64_t)v * ldv] * y[v];
        x[i] += s;
    }
}

__global__ void k_scale_copy(int64_t n, double a, const double *__restrict__ x, double *__restrict__ y) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = a * x[i];
}

__global__ void k_zero_d(int64_t n, double *__restrict__ y) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = 0.0;
}

unsigned vgrid(int64_t n) {
    int64_t g = omx_cdiv(n, 256);
    return (unsigned)(g > 16384 ? 16384 : (g < 1 ? 1 : g));
}

int dots(omx_solver *s, int64_t n, int nv, const double *V, int64_t ldv, const double *w, double *out,
         hipStream_t st) {
    double *part = static_cast<double *>(s->dots.p);
    hipLaunchKernelGGL(k_dot_partial, dim3(kDotBlocks), dim3(256), 0, st, n, nv, V, ldv, w, part);
    OMX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_dot_final, dim3((unsigned)omx_cdiv(nv, 64)), dim3(64), 0, st, nv, kDotBlocks, part, out);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

double wall_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// GMRES basis (restart + 1 vectors of n), Hessenberg column and dot-product
// partials for restarts up to `restart`; allocates only when the capacity
// grows (omx_solver_create, omx_solver_reserve_gmres).
int reserve_gmres(omx_solver *s, int restart) {
    if (restart <= s->restart_cap) return OMX_OK;
    dfree(s->V);
    dfree(s->hv);
    dfree(s->dots);
    s->restart_cap = 0;
    RC(dalloc(s->V, (size_t)(restart + 1) * s->n * 8));
    RC(dalloc(s->hv, (size_t)(restart + 2) * 8));
    RC(dalloc(s->dots, (size_t)(restart + 2) * kDotBlocks * 8));
    s->restart_cap = restart;
    return OMX_OK;
}

}  // namespace

extern "C" {

int omx_solver_create(int64_t n, int64_t b, int alt, int flags, omx_solver **out) {
    if (!out || n <= 0 || b <= 0 || n % b || b % OMX_STRIP || alt < 0 || alt > 2) return OMX_ERR_ARG;
    omx_solver *s = new omx_solver();
    s->n = n;
    s->b = b;
    s->nb = n / b;
    s->alt = alt;
    s->fp8 = (flags & OMX_SOLVER_FP8) ? 1 : 0;
    s->scale16 = (flags & OMX_SOLVER_SCALE16) ? 1 : 0;
    s->G = (flags & OMX_SOLVER_GROUP4) ? 4 : 2;
    s->timers.on = (flags & OMX_SOLVER_TIMERS) != 0;
    int rc = OMX_OK;
    auto fail = [&](int code) {
        omx_solver_destroy(s);
        return code;
    };
    if (hipStreamCreateWithFlags(&s->crit, hipStreamNonBlocking) != hipSuccess) return fail(OMX_ERR_HIP);
    if (hipStreamCreateWithFlags(&s->bulk, hipStreamNonBlocking) != hipSuccess) return fail(OMX_ERR_HIP);
    const int64_t nb = s->nb;
    s->ev_la.resize(s->G);
    for (auto &e : s->ev_la)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return fail(OMX_ERR_HIP);
    s->ev_panel.resize(nb);
    s->ev_t0.resize(nb);
    s->ev_t1.resize(nb);
    if (hipEventCreate(&s->ev_start) != hipSuccess) return fail(OMX_ERR_HIP);
    s->ev_grp.resize(nb / s->G + 2);
    for (auto &e : s->ev_grp)
        if (hipEventCreate(&e) != hipSuccess) return fail(OMX_ERR_HIP);
    for (int64_t k = 0; k < nb; ++k) {
        if (hipEventCreateWithFlags(&s->ev_panel[k], hipEventDisableTiming) != hipSuccess) return fail(OMX_ERR_HIP);
        if (hipEventCreate(&s->ev_t0[k]) != hipSuccess) return fail(OMX_ERR_HIP);
        if (hipEventCreate(&s->ev_t1[k]) != hipSuccess) return fail(OMX_ERR_HIP);
    }
    const int64_t rows = (n > b) ? n - b : b;
    // group buffers: packed rows from the group's first panel row, K = G b
    const int64_t kg = s->G * b;
    const size_t pb16 = omx_panel_bytes(rows, kg, 0);
    for (int i = 0; i < 2; ++i) {
        if ((rc = dalloc(s->Lp[i], pb16)) || (rc = dalloc(s->Rp[i], pb16))) return fail(rc);
        if (s->fp8) {
            const size_t pb8 = omx_panel_bytes(rows, kg, 1);
            const size_t ne = omx_panel_exps_count(rows, kg);
            if ((rc = dalloc(s->L8[i], pb8)) || (rc = dalloc(s->R8[i], pb8)) ||
                (rc = dalloc(s->Le[i], ne * 4)) || (rc = dalloc(s->Re[i], ne * 4)))
                return fail(rc);
        }
    }
    if (s->scale16)
        for (int i = 0; i < 2 * s->G; ++i)
            if ((rc = dalloc(s->Lse[i], omx_strips(rows) * 4)) || (rc = dalloc(s->Rse[i], omx_strips(rows) * 4)))
                return fail(rc);
    if (s->fp8 || s->scale16) {
        const size_t ne = omx_panel_exps_count(rows, b);
        if ((rc = dalloc(s->L8max, ne * 4)) || (rc = dalloc(s->R8max, ne * 4))) return fail(rc);
    }
    if (s->fp8) {
        if ((rc = dalloc(s->pol, OMX_MLP_NPARAMS * 4)) || (rc = dalloc(s->feat, (size_t)nb * OMX_MLP_NFEATURES * 4)) ||
            (rc = dalloc(s->mlp_h, 2 * 64 * 4)) || (rc = dalloc(s->zdev, 16)))
            return fail(rc);
        if (hipEventCreateWithFlags(&s->ev_dec, hipEventDisableTiming) != hipSuccess) return fail(OMX_ERR_HIP);
        if (hipHostMalloc(reinterpret_cast<void **>(&s->zhost), 16) != hipSuccess) return fail(OMX_ERR_HIP);
    }
    s->last_sched.assign(nb, 0);
    s->prec.assign(nb, 0);
    s->last_feat.assign((size_t)nb * OMX_MLP_NFEATURES, 0.0f);
    if ((rc = dalloc(s->Tp, omx_panel_bytes(rows, b, 0)))) return fail(rc);
    if ((rc = dalloc(s->Wup, (size_t)b * b * 4)) || (rc = dalloc(s->Wlo, (size_t)b * b * 4)) ||
        (rc = dalloc(s->Wtmp, (size_t)b * b * 4)) ||
        (rc = dalloc(s->Wup_h, omx_panel_bytes(b, b, 0))) || (rc = dalloc(s->Wlo_h, omx_panel_bytes(b, b, 0))))
        return fail(rc);
    if ((rc = dalloc(s->info, (size_t)nb * sizeof(int) + 16))) return fail(rc);
    // refinement vectors
    if ((rc = dalloc(s->r, n * 8)) || (rc = dalloc(s->w, n * 8)) ||
        (rc = dalloc(s->gemv_ws, std::max(std::max<size_t>(omx_otf_gemv_workspace_bytes(n), 8), omx_lowrank_gemv_workspace_bytes(16)))) ||
        (rc = dalloc(s->trsv_ws, omx_trsv_workspace_bytes(n))) || (rc = dalloc(s->nrm, 64)))
        return fail(rc);
    if ((rc = reserve_gmres(s, OMX_GMRES_RESTART_DEFAULT))) return fail(rc);
    *out = s;
    return OMX_OK;
}

int omx_solver_reserve_gmres(omx_solver *s, int restart) {
    if (!s || restart <= 0 || restart > OMX_GMRES_RESTART_MAX) return OMX_ERR_ARG;
    return reserve_gmres(s, restart);
}

int omx_solver_destroy(omx_solver *s) {
    if (!s) return OMX_OK;
    if (s->crit) (void)hipStreamSynchronize(s->crit);
    if (s->bulk) (void)hipStreamSynchronize(s->bulk);
    if (s->ev_start) (void)hipEventDestroy(s->ev_start);
    for (auto *v : {&s->ev_panel, &s->ev_t0, &s->ev_t1, &s->ev_la, &s->ev_grp})
        for (auto e : *v)
            if (e) (void)hipEventDestroy(e);
    DevBuf *bufs[] = {&s->Lp[0], &s->Lp[1], &s->Rp[0], &s->Rp[1], &s->Tp, &s->L8[0], &s->L8[1], &s->R8[0],
                      &s->R8[1], &s->Le[0], &s->Le[1], &s->Re[0], &s->Re[1], &s->L8max, &s->R8max,
                      &s->Wup, &s->Wlo, &s->Wup_h, &s->Wlo_h, &s->Wtmp, &s->info, &s->r,
                      &s->w, &s->gemv_ws, &s->trsv_ws, &s->nrm, &s->V, &s->hv, &s->dots,
                      &s->pol, &s->feat, &s->mlp_h, &s->zdev};
    for (DevBuf *d : bufs) dfree(*d);
    for (int i = 0; i < 8; ++i) {
        dfree(s->Lse[i]);
        dfree(s->Rse[i]);
    }
    if (s->ev_dec) (void)hipEventDestroy(s->ev_dec);
    for (auto e : s->timers.pool) (void)hipEventDestroy(e);
    if (s->zhost) (void)hipHostFree(s->zhost);
    if (s->crit) (void)hipStreamDestroy(s->crit);
    if (s->bulk) (void)hipStreamDestroy(s->bulk);
    delete s;
    return OMX_OK;
}

// Packed coordinates of panel i in groups of G: group buffer (i/G) % 2,
// K window (i % G) b, packed row 0 = matrix row (G (i/G) + 1) b, the first
// row of the group's first column panel.
static inline int pbuf(int64_t i, int64_t G) { return (int)((i / G) & 1); }
static inline int64_t porigin(int64_t i, int64_t b, int64_t G) { return ((i / G) * G + 1) * b; }
static inline int64_t pwin(int64_t i, int64_t b, int64_t G) { return (i % G) * b; }

// A(r0:r0+m, c0:c0+nn) -= L R^T over K window [k0, k0 + k) of panel i's
// group buffer (the panels whose windows it spans, from panel i on).
static int schur_window(omx_solver *s, int64_t i, int64_t k, int64_t k0, int64_t r0, int64_t c0, int64_t m,
                        int64_t nn, float *A, int64_t lda, hipStream_t st) {
    const int64_t b = s->b, G = s->G, ldk = G * b;
    const int buf = pbuf(i, G);
    const int64_t xr0 = r0 - porigin(i, b, G), yr0 = c0 - porigin(i, b, G);
    float *C = A + r0 + c0 * lda;
    if (s->prec[i])
        return omx_gemm_f8_packed_window(m, nn, k, -1.0f, static_cast<uint8_t *>(s->L8[buf].p),
                                         static_cast<int32_t *>(s->Le[buf].p), xr0, ldk, k0,
                                         static_cast<uint8_t *>(s->R8[buf].p), static_cast<int32_t *>(s->Re[buf].p),
                                         yr0, ldk, k0, 1.0f, C, lda, st);
    const int32_t *xe = nullptr, *ye = nullptr;
    if (s->scale16) {
        xe = static_cast<int32_t *>(s->Lse[G * buf + i % G].p);
        ye = static_cast<int32_t *>(s->Rse[G * buf + i % G].p);
    }
    return omx_gemm_f16_packed_window(m, nn, k, -1.0f, static_cast<uint16_t *>(s->Lp[buf].p), xe, xr0, ldk, k0,
                                      static_cast<uint16_t *>(s->Rp[buf].p), ye, yr0, ldk, k0, 1.0f, C, lda, st);
}

// A(r0.., c0..) (m x nn) -= the panels i0 .. i1 (all in one group), in
// order; each run of consecutive panels packed alike is one product over
// its K windows (scaled FP16 panels, whose scales are per panel, one by one).
static int prec_class(int p) { return p > 0 ? 1 : (p < 0 ? -1 : 0); }

static int apply_panels(omx_solver *s, int64_t i0, int64_t i1, int64_t r0, int64_t c0, int64_t m, int64_t nn,
                        float *A, int64_t lda, hipStream_t st) {
    if (m <= 0 || nn <= 0) return OMX_OK;
    const int64_t b = s->b, G = s->G;
    for (int64_t i = i0; i <= i1;) {
        int64_t j = i;
        const int cls = prec_class(s->prec[i]);
        if (cls != 0 || !s->scale16)
            while (j + 1 <= i1 && (j + 1) / G == i / G && prec_class(s->prec[j + 1]) == cls) ++j;
        if (cls < 0) {
            // FP32 class: C -= A(r0.., ib..jb+b) A(ib..jb+b, c0..), the
            // panels as factored (column panels side by side, row panels
            // stacked), on the FP32 matrix cores
            const int64_t ib = i * b, k = (j - i + 1) * b;
            RC(omx_sgemm_f32(m, nn, k, -1.0f, A + r0 + ib * lda, lda, A + ib + c0 * lda, lda, 1.0f,
                             A + r0 + c0 * lda, lda, st));
        } else {
            RC(schur_window(s, i, (j - i + 1) * b, pwin(i, b, G), r0, c0, m, nn, A, lda, st));
        }
        i = j + 1;
    }
    return OMX_OK;
}

// Block row/col t (A_tt, A[t+1:, t], A[t, t+1:]) -= panels i0 .. i1.
static int update_block(omx_solver *s, int64_t t, int64_t i0, int64_t i1, float *A, int64_t lda, hipStream_t st) {
    if (s->skip & OMX_SKIP_GEMM) return OMX_OK;
    const int64_t n = s->n, b = s->b, tb = t * b;
    const double fl = 2.0 * (double)(i1 - i0 + 1) * b * b * (2.0 * (n - tb) - b);
    long sp;
    RC(s->timers.beg(kT_GEMM_UPDATE, st, &sp));
    RC(apply_panels(s, i0, i1, tb, tb, n - tb, b, A, lda, st));
    RC(apply_panels(s, i0, i1, tb, tb + b, b, n - tb - b, A, lda, st));
    RC(s->timers.end(sp, st, fl));
    s->crit_flops += fl;
    return OMX_OK;
}

// Pack panel i (its finished FP32 column / row panels) into its group
// buffer window, in the precision s->prec[i].
static int pack_panel(omx_solver *s, int64_t i, int have_max, int64_t rest, const float *Acol, const float *Arow,
                      int64_t lda, hipStream_t st) {
    if (s->prec[i] < 0) return OMX_OK;  // FP32 class: the products read the panels in A
    const int64_t b = s->b, G = s->G, ldk = G * b;
    const int buf = pbuf(i, G);
    const int64_t rstrip = pwin(i, b, G) / OMX_STRIP;  // packed rows start at (i % G) b
    const int64_t k0 = pwin(i, b, G);
    if (s->prec[i]) {
        const int h = s->prec[i] - 1;
        const size_t sb8 = (size_t)(ldk / OMX_KSTAGE8) * 32768, ne = (size_t)(ldk / OMX_SCALE_TILE);
        uint8_t *L8 = static_cast<uint8_t *>(s->L8[buf].p) + rstrip * sb8;
        uint8_t *R8 = static_cast<uint8_t *>(s->R8[buf].p) + rstrip * sb8;
        int32_t *Le = static_cast<int32_t *>(s->Le[buf].p) + rstrip * ne;
        int32_t *Re = static_cast<int32_t *>(s->Re[buf].p) + rstrip * ne;
        TIMED(s->timers, kT_CONV_L, st, (double)rest * b,
              omx_pack_f8_window(rest, b, Acol, 1, lda, h, have_max, L8, Le, ldk, k0,
                                 static_cast<float *>(s->L8max.p), st));
        TIMED(s->timers, kT_CONV_R, st, (double)rest * b,
              omx_pack_f8_window(rest, b, Arow, lda, 1, h, have_max, R8, Re, ldk, k0,
                                 static_cast<float *>(s->R8max.p), st));
        return OMX_OK;
    }
    const size_t sb16 = (size_t)(ldk / OMX_KSTAGE16) * 16384;  // uint16 per strip
    uint16_t *Lp = static_cast<uint16_t *>(s->Lp[buf].p) + rstrip * sb16;
    uint16_t *Rp = static_cast<uint16_t *>(s->Rp[buf].p) + rstrip * sb16;
    int32_t *xe = nullptr, *ye = nullptr;
    if (s->scale16) {
        xe = static_cast<int32_t *>(s->Lse[G * buf + i % G].p) + rstrip;
        ye = static_cast<int32_t *>(s->Rse[G * buf + i % G].p) + rstrip;
    }
    TIMED(s->timers, kT_CONV_L, st, (double)rest * b,
          omx_pack_f16_window(rest, b, Acol, 1, lda, have_max, Lp, xe, static_cast<float *>(s->L8max.p), ldk, k0, st));
    TIMED(s->timers, kT_CONV_R, st, (double)rest * b,
          omx_pack_f16_window(rest, b, Arow, lda, 1, have_max, Rp, ye, static_cast<float *>(s->R8max.p), ldk, k0, st));
    return OMX_OK;
}

// Statistics of step k's finished panels (FP8 solvers): tile maxima of the
// column and row panels (unless the panel products left them: have_max)
// and the 16 scheduler features; with a policy, the MLP's outputs are
// copied to pinned host memory behind s->ev_dec.
static int step_features(omx_solver *s, int64_t k, int64_t rest, const float *Akk, const float *Acol,
                         const float *Arow, int64_t lda, int prev, bool infer, bool have_max, hipStream_t cs) {
    const int64_t b = s->b;
    float *lmax = static_cast<float *>(s->L8max.p), *rmax = static_cast<float *>(s->R8max.p);
    if (!have_max) {
        RC(omx_panel_tile_maxabs(rest, b, Acol, 1, lda, lmax, cs));
        RC(omx_panel_tile_maxabs(rest, b, Arow, lda, 1, rmax, cs));
    }
    const int64_t nt = (int64_t)omx_panel_exps_count(rest, b);
    float *f = static_cast<float *>(s->feat.p) + k * OMX_MLP_NFEATURES;
    RC(omx_sched_features_f32(nt, lmax, nt, rmax, Akk, lda, b, k, s->nb, s->n, prev, f, cs));
    if (infer) {
        float *h = static_cast<float *>(s->mlp_h.p);
        float *z = static_cast<float *>(s->zdev.p);
        RC(omx_mlp_forward_f32(1, static_cast<const float *>(s->pol.p), f, h, h + 64, z, cs));
        OMX_HIP_CHECK(hipMemcpyAsync(s->zhost, z, 3 * sizeof(float), hipMemcpyDeviceToHost, cs));
        OMX_HIP_CHECK(hipEventRecord(s->ev_dec, cs));
    }
    return OMX_OK;
}

// The policy's choice for the step whose outputs step_features() queued:
// 0 FP16, 1 + headroom for FP8.
static int policy_decision(omx_solver *s, int *choice) {
    OMX_HIP_CHECK(hipEventSynchronize(s->ev_dec));
    const float *z = s->zhost;
    if (!(z[1] > z[0])) {
        *choice = 0;
        return OMX_OK;
    }
    float h = std::nearbyint(z[2]);
    if (!(h >= 0.0f)) h = 0.0f;
    if (h > 3.0f) h = 3.0f;
    *choice = 1 + (int)h;
    return OMX_OK;
}

int omx_solver_set_skip(omx_solver *s, int mask) {
    if (!s || mask < 0 || mask > 15) return OMX_ERR_ARG;
    s->skip = mask & (OMX_SKIP_GETRF | OMX_SKIP_TRSM | OMX_SKIP_GEMM);
    return OMX_OK;
}

int omx_solver_step_times(omx_solver *s, double *ms) {
    if (!s || !ms || s->ev_grp.empty()) return OMX_ERR_ARG;
    const int64_t G = s->G;
    for (int64_t k = 0; k < s->nb; ++k) {
        const int64_t g = std::min(k / G, s->ngrp_done);
        float t = 0.0f;
        OMX_HIP_CHECK(hipEventElapsedTime(&t, s->ev_start, s->ev_grp[g]));
        ms[k] = t;
    }
    return OMX_OK;
}

int omx_solver_set_policy(omx_solver *s, const float *params) {
    if (!s) return OMX_ERR_ARG;
    if (!params) {
        s->has_policy = 0;
        return OMX_OK;
    }
    if (!s->fp8) return OMX_ERR_ARG;  // the policy needs the FP8 buffers
    OMX_HIP_CHECK(hipMemcpy(s->pol.p, params, OMX_MLP_NPARAMS * sizeof(float), hipMemcpyHostToDevice));
    s->has_policy = 1;
    return OMX_OK;
}

int omx_solver_last_schedule(omx_solver *s, int *sched, float *features) {
    if (!s) return OMX_ERR_ARG;
    if (sched) std::memcpy(sched, s->last_sched.data(), s->last_sched.size() * sizeof(int));
    if (features) {
        if (!s->fp8) return OMX_ERR_ARG;
        OMX_HIP_CHECK(hipMemcpy(features, s->feat.p, s->last_feat.size() * sizeof(float), hipMemcpyDeviceToHost));
    }
    return OMX_OK;
}

int omx_lu_factor(omx_solver *s, float *A, int64_t lda, const int *sched, double *timings) {
    if (!s || !A || lda < s->n || lda % 4) return OMX_ERR_ARG;
    const int64_t n = s->n, b = s->b, nb = s->nb;
    hipStream_t cs = s->crit, bs = s->bulk;
    if (sched)
        for (int64_t k = 0; k < nb; ++k)
            if (sched[k] < -1 || sched[k] > 4 || (sched[k] > 0 && !s->fp8)) return OMX_ERR_ARG;
    // sched given: fixed choices; else the policy if one is set; else FP16
    const bool learned = !sched && s->has_policy;
    const bool stats = s->fp8 != 0;
    int *info = static_cast<int *>(s->info.p);  // one slot per diagonal block
    uint16_t *Tp = static_cast<uint16_t *>(s->Tp.p);
    float *Wup = static_cast<float *>(s->Wup.p), *Wlo = static_cast<float *>(s->Wlo.p);
    uint16_t *Wup_h = static_cast<uint16_t *>(s->Wup_h.p), *Wlo_h = static_cast<uint16_t *>(s->Wlo_h.p);
    s->bulk_count = 0;
    s->bulk_flops = 0;
    s->crit_flops = 0;
    std::fill(s->prec.begin(), s->prec.end(), 0);
    OMX_HIP_CHECK(hipDeviceSynchronize());
    s->timers.reset();
    PhaseTimers &T = s->timers;
    const double t0 = wall_ms();
    OMX_HIP_CHECK(hipEventRecord(s->ev_start, cs));
    s->ngrp_done = 0;
    if (stats) OMX_HIP_CHECK(hipMemsetAsync(s->feat.p, 0, s->feat.bytes, cs));

    // Factor panel i on the crit stream: getrf of the (up to date) diagonal
    // block, the column and row panels (FP32 TRSM at step 0 and with -alt 0,
    // panel_trf.hpp:188-246; the FP16 inverse products with -alt 1/2,
    // panel_trf_gdirect.hpp:73-183), the step's statistics and precision,
    // and the pack into the pair buffer.
    auto factor = [&](int64_t i) -> int {
        const int64_t ib = i * b, rest = n - ib - b;
        float *Akk = A + ib + ib * lda;
        if (!(s->skip & OMX_SKIP_GETRF))
            TIMED(T, kT_DIAG_LU, cs, 2.0 / 3.0 * b * (double)b * b, omx_getrf_nopiv_f32(b, Akk, lda, info + i, cs));
        int choice = 0;
        if (rest > 0) {
            bool have_max = false;  // the panels' tile maxima already computed
            float *Acol = A + (ib + b) + ib * lda;
            float *Arow = A + ib + (ib + b) * lda;
            const double rb2 = (double)rest * b * b;
            if (s->skip & OMX_SKIP_TRSM) {
                // panels left as they are (the reference's skip bit 2)
            } else if (i == 0 || s->alt == 0 || (sched && sched[i] < 0)) {
                // FP32 panels by TRSM: step 0 (the reference's), -alt 0, and
                // the FP32 Schur class (whose products take the panels as
                // they are, so the FP16 inverse products would limit them)
                if (i == 0) {
                    // nothing is queued on the bulk stream before the first
                    // group's panels: the two independent solves run side by
                    // side (the two look-ahead events borrowed here are
                    // recorded afresh before anything waits on them again)
                    OMX_HIP_CHECK(hipEventRecord(s->ev_la[0], cs));
                    OMX_HIP_CHECK(hipStreamWaitEvent(bs, s->ev_la[0], 0));
                    TIMED(T, kT_TRSM_R, bs, rb2, omx_trsm_f32(OMX_TRSM_LLNU, b, rest, Akk, lda, Arow, lda, bs));
                    TIMED(T, kT_TRSM_L, cs, rb2, omx_trsm_f32(OMX_TRSM_RUNN, rest, b, Akk, lda, Acol, lda, cs));
                    OMX_HIP_CHECK(hipEventRecord(s->ev_la[1], bs));
                    OMX_HIP_CHECK(hipStreamWaitEvent(cs, s->ev_la[1], 0));
                } else {
                    TIMED(T, kT_TRSM_L, cs, rb2, omx_trsm_f32(OMX_TRSM_RUNN, rest, b, Akk, lda, Acol, lda, cs));
                    TIMED(T, kT_TRSM_R, cs, rb2, omx_trsm_f32(OMX_TRSM_LLNU, b, rest, Akk, lda, Arow, lda, cs));
                }
            } else {
                // U^-1 and L^-1 of the factored diagonal block
                TIMED(T, kT_DIAG_LU, cs, 2.0 / 3.0 * b * (double)b * b,
                      omx_trtri_lu_f32(b, Akk, lda, Wup, Wlo, b, static_cast<float *>(s->Wtmp.p), cs));
                // Y(j, l) = W(l, j), upper part only for -alt 2
                long sp;
                RC(T.beg(kT_CONV_L, cs, &sp));
                RC(omx_pack_f16(b, b, Wup, b, 1, s->alt == 2 ? OMX_MASK_K_LE_R : OMX_MASK_NONE, Wup_h, nullptr, cs));
                RC(omx_pack_f16(rest, b, Acol, 1, lda, OMX_MASK_NONE, Tp, nullptr, cs));
                RC(T.end(sp, cs, (double)(rest + b) * b));
                // U^-1 is upper triangular: each output column tile stops
                // at its diagonal k tile (the rest of its products are
                // zero).  With the statistics on, the product also leaves
                // the column panel's tile maxima (no separate pass over it).
                float *lmax = stats ? static_cast<float *>(s->L8max.p) : nullptr;
                float *rmax = stats ? static_cast<float *>(s->R8max.p) : nullptr;
                TIMED(T, kT_TRSM_L, cs, 2.0 * rb2,
                      omx_gemm_f16_packed_tri(rest, b, b, 1.0f, Tp, Wup_h, 1, 0.0f, Acol, lda, lmax, cs));
                // X(i, l) = W(i, l), strict lower + unit diagonal for -alt 2
                RC(T.beg(kT_CONV_R, cs, &sp));
                RC(omx_pack_f16(b, b, Wlo, 1, b, s->alt == 2 ? OMX_MASK_R_GT_K_UNIT : OMX_MASK_NONE, Wlo_h, nullptr,
                                cs));
                RC(omx_pack_f16(rest, b, Arow, lda, 1, OMX_MASK_NONE, Tp, nullptr, cs));
                RC(T.end(sp, cs, (double)(rest + b) * b));
                // L^-1 lower triangular: each output row tile likewise
                TIMED(T, kT_TRSM_R, cs, 2.0 * rb2,
                      omx_gemm_f16_packed_tri(b, rest, b, 1.0f, Wlo_h, Tp, 2, 0.0f, Arow, lda, rmax, cs));
                have_max = stats;
            }
            s->crit_flops += 4.0 * rest * (double)b * b;
            const int prev = (i > 0 && s->prec[i - 1] > 0) ? 1 : 0;
            if (stats) RC(step_features(s, i, rest, Akk, Acol, Arow, lda, prev, learned, have_max, cs));
            // with a policy the host waits here for the step's outputs; the
            // current pair's trailing update is already queued on the bulk
            // stream, so the GPU stays busy
            if (learned) RC(policy_decision(s, &choice));
            else if (sched) choice = sched[i];
            s->prec[i] = choice;
            RC(pack_panel(s, i, stats, rest, Acol, Arow, lda, cs));
        }
        s->prec[i] = choice;
        s->last_sched[i] = choice;
        return OMX_OK;
    };

    // group 0's panels
    const int64_t G = s->G;
    for (int64_t t = 0; t < G && t < nb; ++t) {
        if (t > 0) RC(update_block(s, t, 0, t - 1, A, lda, cs));
        RC(factor(t));
    }
    OMX_HIP_CHECK(hipEventRecord(s->ev_panel[0], cs));
    for (int64_t g = 0; G * (g + 1) < nb; ++g) {
        const int64_t i0 = G * g, i1 = std::min(G * g + G - 1, nb - 1);
        const int64_t h0 = G * (g + 1), h1 = std::min(h0 + G - 1, nb - 1);
        // ---- bulk: the next group's blocks, then the trailing update ----
        OMX_HIP_CHECK(hipStreamWaitEvent(bs, s->ev_panel[g], 0));
        for (int64_t t = h0; t <= h1; ++t) {
            RC(update_block(s, t, i0, i1, A, lda, bs));
            OMX_HIP_CHECK(hipEventRecord(s->ev_la[t - h0], bs));
        }
        const int64_t f = G * (g + 2);
        if (f < nb && !(s->skip & OMX_SKIP_GEMM)) {
            const int64_t m = n - f * b;
            OMX_HIP_CHECK(hipEventRecord(s->ev_t0[s->bulk_count], bs));
            TIMED(T, kT_GEMM_UPDATE, bs, 2.0 * (double)m * m * (double)((i1 - i0 + 1) * b),
                  apply_panels(s, i0, i1, f * b, f * b, m, m, A, lda, bs));
            OMX_HIP_CHECK(hipEventRecord(s->ev_t1[s->bulk_count], bs));
            s->bulk_flops += 2.0 * (double)m * m * (double)((i1 - i0 + 1) * b);
            ++s->bulk_count;
        }
        // every Schur product of group g's panels is done
        OMX_HIP_CHECK(hipEventRecord(s->ev_grp[g], bs));
        s->ngrp_done = g + 1;
        // ---- crit: the next group's panels ----
        // block t holds all earlier groups' updates once its look-ahead
        // products are done; the later blocks' products still run on the
        // bulk stream while block t is factored
        for (int64_t t = h0; t <= h1; ++t) {
            OMX_HIP_CHECK(hipStreamWaitEvent(cs, s->ev_la[t - h0], 0));
            if (t > h0) RC(update_block(s, t, h0, t - 1, A, lda, cs));
            RC(factor(t));
        }
        OMX_HIP_CHECK(hipEventRecord(s->ev_panel[g + 1], cs));
    }
    // the last group's panels update only blocks of their own group (crit)
    OMX_HIP_CHECK(hipEventRecord(s->ev_grp[s->ngrp_done], cs));
    OMX_HIP_CHECK(hipStreamSynchronize(cs));
    OMX_HIP_CHECK(hipStreamSynchronize(bs));
    s->t_lu_ms = wall_ms() - t0;
    RC(T.collect());
    T.seconds[kT_TOTAL] += s->t_lu_ms * 1e-3;
    s->bulk_ms = 0;
    for (int i = 0; i < s->bulk_count; ++i) {
        float ms = 0;
        OMX_HIP_CHECK(hipEventElapsedTime(&ms, s->ev_t0[i], s->ev_t1[i]));
        s->bulk_ms += ms;
    }
    std::vector<int> hinfo(nb, 0);
    OMX_HIP_CHECK(hipMemcpy(hinfo.data(), info, sizeof(int) * nb, hipMemcpyDeviceToHost));
    if (timings) {
        timings[0] = s->t_lu_ms;
        timings[1] = s->bulk_ms;
        timings[2] = s->bulk_count;
        timings[3] = s->bulk_flops;
        timings[4] = s->crit_flops;
    }
    if (!(s->skip & OMX_SKIP_GETRF))
        for (int64_t i = 0; i < nb; ++i)
            if (hinfo[i] != 0x7fffffff) return (int)(i * b + hinfo[i]);
    return OMX_OK;
}

// Refinement.  method 0: classical IR (iterative_refinement.hpp:141-171);
// stop_harness = 16 is the reference's test; stop_harness <= 0 refines for
// exactly maxit residual evaluations (the FP64 "refinement to convergence"
// used as the large-n forward-error oracle).  x_given = 1 starts from the x
// passed in instead of rhs./diag.
// method 1: GMRES-IR (same outer loop; the correction solves A d = r by
// GMRES(restart) on (LU)^-1 A d = (LU)^-1 r with classical Gram-Schmidt
// applied twice, all in FP64).  Returns OMX_OK when the harness fell below
// 16, OMX_ERR_NOCONV otherwise.  out[0..3] = residual, harness, corrections,
// total inner GMRES iterations; trace (2*maxit) = residual, harness per step.
int omx_ir_solve(omx_solver *s, const float *LU, int64_t lda, uint64_t seed, const double *diag,
                 const double *rhs, double *x, int method, int maxit, int restart, double inner_tol,
                 double stop_harness, int x_given, double *out, double *trace) {
    omx_system sys{};
    sys.kind = OMX_SYSTEM_LCG;
    sys.seed = seed;
    sys.diag = diag;
    return omx_ir_solve_system(s, LU, lda, &sys, rhs, x, method, maxit, restart, inner_tol, stop_harness,
                               x_given, out, trace);
}

// y += alpha A x for the system being refined
static int system_gemv(omx_solver *s, const omx_system *sys, double alpha, const double *x, double *y,
                       hipStream_t st) {
    if (sys->kind == OMX_SYSTEM_LCG)
        return omx_otf_gemv_f64(s->n, sys->seed, sys->diag, alpha, x, y, s->gemv_ws.p, s->gemv_ws.bytes, st);
    if (sys->kind == OMX_SYSTEM_HIGHAM) return omx_higham_gemv_f64(s->n, sys->alpha, sys->beta, alpha, x, y, st);
    return omx_lowrank_gemv_f64(s->n, sys->rank, sys->sigma, sys->P, sys->Q, alpha, x, y, s->gemv_ws.p,
                                s->gemv_ws.bytes, st);
}

// r <- U^-1 L^-1 r with the FP32 factors (one IR_TRSV phase)
static int trsv_pair(omx_solver *s, const float *LU, int64_t lda, double *r, hipStream_t st) {
    const int64_t n = s->n, b = s->b;
    long sp;
    RC(s->timers.beg(kT_IR_TRSV, st, &sp));
    RC(omx_trsv_lower_unit_f32f64(n, b, LU, lda, r, s->trsv_ws.p, s->trsv_ws.bytes, st));
    RC(omx_trsv_upper_f32f64(n, b, LU, lda, r, s->trsv_ws.p, s->trsv_ws.bytes, st));
    return s->timers.end(sp, st, 2.0 * n * (double)n);
}

static int ir_solve_impl(omx_solver *s, const float *LU, int64_t lda, const omx_system *sys, const double *rhs,
                         double *x, int method, int maxit, int restart, double inner_tol, double stop_harness,
                         int x_given, double *out, double *trace) {
    if (!s || !LU || !sys || !sys->diag || !rhs || !x || maxit <= 0 || lda < s->n) return OMX_ERR_ARG;
    const bool lowrank_ok = sys->kind == OMX_SYSTEM_LOWRANK && sys->rank >= 0 && sys->rank <= 16 && sys->sigma &&
                            (sys->rank == 0 || (sys->P && sys->Q)) && sys->norm_inf > 0;
    const bool higham_ok = sys->kind == OMX_SYSTEM_HIGHAM && std::isfinite(sys->alpha) &&
                           std::isfinite(sys->beta) && sys->norm_inf > 0;
    if (sys->kind != OMX_SYSTEM_LCG && !lowrank_ok && !higham_ok) return OMX_ERR_ARG;
    const double *diag = sys->diag;
    // only the two methods exist; GMRES's basis was sized by
    // omx_solver_create (restart 50) or omx_solver_reserve_gmres: this call
    // allocates nothing
    if (method != 0 && method != 1) return OMX_ERR_ARG;
    if (method == 1 && (restart <= 0 || restart > OMX_GMRES_RESTART_MAX)) return OMX_ERR_ARG;
    if (method == 1 && restart > s->restart_cap) return OMX_ERR_WORKSPACE;
    const int64_t n = s->n, b = s->b;
    hipStream_t st = s->crit;
    double *r = static_cast<double *>(s->r.p), *w = static_cast<double *>(s->w.p);
    double *nrm = static_cast<double *>(s->nrm.p);
    double hn[4];
    RC(omx_norm_inf_f64(n, diag, nrm, st));
    RC(omx_norm_inf_f64(n, rhs, nrm + 1, st));
    OMX_HIP_CHECK(hipMemcpyAsync(hn, nrm, 16, hipMemcpyDeviceToHost, st));
    OMX_HIP_CHECK(hipStreamSynchronize(st));
    const double norma = sys->kind == OMX_SYSTEM_LCG ? 2.0 * hn[0] : sys->norm_inf, normb = hn[1];
    if (!x_given) RC(omx_div_f64(n, rhs, diag, x, st));
    int inner_total = 0, result = OMX_ERR_NOCONV, it = 0;
    double normr = -1, harness = -1;
    for (it = 0; it < maxit; ++it) {
        RC(omx_copy_f64(n, rhs, r, st));
        RC(omx_norm_inf_f64(n, x, nrm, st));
        TIMED(s->timers, kT_IR_GEMV, st, 2.0 * n * (double)n, system_gemv(s, sys, -1.0, x, r, st));
        RC(omx_norm_inf_f64(n, r, nrm + 1, st));
        OMX_HIP_CHECK(hipMemcpyAsync(hn, nrm, 16, hipMemcpyDeviceToHost, st));
        OMX_HIP_CHECK(hipStreamSynchronize(st));
        const double normx = hn[0];
        normr = hn[1];
        harness = omx_harness(normr, norma, normx, normb, n);
        if (trace) {
            trace[2 * it] = normr;
            trace[2 * it + 1] = harness;
        }
        if (harness < stop_harness) {
            result = OMX_OK;
            break;
        }
        if (!std::isfinite(harness)) break;
        if (it == maxit - 1) {  // refined to the iteration cap without a stop test
            if (stop_harness <= 0.0) result = OMX_OK;
            break;
        }
        if (method == 0) {
            RC(trsv_pair(s, LU, lda, r, st));
            RC(omx_axpy_f64(n, 1.0, r, x, st));
            continue;
        }
        // ---- GMRES(restart) on the preconditioned correction equation ----
        const int m = restart;
        double *V = static_cast<double *>(s->V.p), *hv = static_cast<double *>(s->hv.p);
        RC(trsv_pair(s, LU, lda, r, st));
        RC(dots(s, n, 1, r, n, r, hv, st));
        double beta = 0;
        OMX_HIP_CHECK(hipMemcpyAsync(&beta, hv, 8, hipMemcpyDeviceToHost, st));
        OMX_HIP_CHECK(hipStreamSynchronize(st));
        beta = std::sqrt(beta);
        if (!(beta > 0) || !std::isfinite(beta)) break;
        hipLaunchKernelGGL(k_scale_copy, dim3(vgrid(n)), dim3(256), 0, st, n, 1.0 / beta, r, V);
        OMX_LAUNCH_CHECK();
        std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), g(m + 1, 0.0), y(m), h(m + 2);
        g[0] = beta;
        int jdone = 0;
        for (int j = 0; j < m; ++j) {
            double *vj = V + (int64_t)j * n;
            hipLaunchKernelGGL(k_zero_d, dim3(vgrid(n)), dim3(256), 0, st, n, w);
            OMX_LAUNCH_CHECK();
            TIMED(s->timers, kT_IR_GEMV, st, 2.0 * n * (double)n, system_gemv(s, sys, 1.0, vj, w, st));
            RC(trsv_pair(s, LU, lda, w, st));
            // classical Gram-Schmidt, twice
            std::fill(h.begin(), h.end(), 0.0);
            for (int pass = 0; pass < 2; ++pass) {
                RC(dots(s, n, j + 1, V, n, w, hv, st));
                hipLaunchKernelGGL(k_multi_axpy, dim3(vgrid(n)), dim3(256), 0, st, n, j + 1, V, n, hv, w);
                OMX_LAUNCH_CHECK();
                std::vector<double> hp(j + 1);
                OMX_HIP_CHECK(hipMemcpyAsync(hp.data(), hv, 8 * (j + 1), hipMemcpyDeviceToHost, st));
                OMX_HIP_CHECK(hipStreamSynchronize(st));
                for (int i = 0; i <= j; ++i) h[i] += hp[i];
            }
            RC(dots(s, n, 1, w, n, w, hv, st));
            double hn2 = 0;
            OMX_HIP_CHECK(hipMemcpyAsync(&hn2, hv, 8, hipMemcpyDeviceToHost, st));
            OMX_HIP_CHECK(hipStreamSynchronize(st));
            const double hnext = std::sqrt(hn2);
            if (hnext > 0) {
                hipLaunchKernelGGL(k_scale_copy, dim3(vgrid(n)), dim3(256), 0, st, n, 1.0 / hnext, w,
                                   V + (int64_t)(j + 1) * n);
                OMX_LAUNCH_CHECK();
            }
            for (int i = 0; i <= j; ++i) H[i + (size_t)j * (m + 1)] = h[i];
            H[(j + 1) + (size_t)j * (m + 1)] = hnext;
            for (int i = 0; i < j; ++i) {
                const double a = H[i + (size_t)j * (m + 1)], bb = H[(i + 1) + (size_t)j * (m + 1)];
                H[i + (size_t)j * (m + 1)] = cs[i] * a + sn[i] * bb;
                H[(i + 1) + (size_t)j * (m + 1)] = -sn[i] * a + cs[i] * bb;
            }
            const double a = H[j + (size_t)j * (m + 1)], bb = H[(j + 1) + (size_t)j * (m + 1)];
            const double den = std::hypot(a, bb);
            cs[j] = den > 0 ? a / den : 1.0;
            sn[j] = den > 0 ? bb / den : 0.0;
            H[j + (size_t)j * (m + 1)] = den;
            H[(j + 1) + (size_t)j * (m + 1)] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            jdone = j + 1;
            ++inner_total;
            if (std::fabs(g[j + 1]) <= inner_tol * beta || hnext == 0.0) break;
        }
        for (int i = jdone - 1; i >= 0; --i) {
            double sum = g[i];
            for (int q = i + 1; q < jdone; ++q) sum -= H[i + (size_t)q * (m + 1)] * y[q];
            y[i] = sum / H[i + (size_t)i * (m + 1)];
        }
        OMX_HIP_CHECK(hipMemcpyAsync(hv, y.data(), 8 * jdone, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_multi_axpy_add, dim3(vgrid(n)), dim3(256), 0, st, n, jdone, V, n, hv, x);
        OMX_LAUNCH_CHECK();
    }
    OMX_HIP_CHECK(hipStreamSynchronize(st));
    if (out) {
        out[0] = normr;
        out[1] = harness;
        out[2] = (result == OMX_OK) ? it : -1;
        out[3] = inner_total;
    }
    return result;
}

int omx_ir_solve_system(omx_solver *s, const float *LU, int64_t lda, const omx_system *sys, const double *rhs,
                        double *x, int method, int maxit, int restart, double inner_tol, double stop_harness,
                        int x_given, double *out, double *trace) {
    if (!s) return OMX_ERR_ARG;
    const double t0 = wall_ms();
    const int rc = ir_solve_impl(s, LU, lda, sys, rhs, x, method, maxit, restart, inner_tol, stop_harness, x_given,
                                 out, trace);
    if (s->timers.on) {
        OMX_HIP_CHECK(hipStreamSynchronize(s->crit));
        RC(s->timers.collect());
        s->timers.seconds[kT_TOTAL] += (wall_ms() - t0) * 1e-3;
    }
    return rc;
}

static const char *const kTimerNames[OMX_TIMER_ITEMS] = {
    "DIAG_BCAST", "LCOL_BCAST", "RROW_BCAST",   "TEST",    "WAIT",         "DIAG_LU", "TRSM_L", "TRSM_R",
    "CONV_L",     "CONV_R",     "GEMM_UPDATE",  "GEMM_PROGRESS", "LAZY_INIT", "WRITE_BACK", "IR_GEMV",
    "IR_GEMV_COMM", "IR_TRSV",  "IR_TRSV_MV",   "IR_TRSV_COMM", "INIT",  "MEMCPY",  "TOTAL",  "MISC"};

int omx_solver_timers(const omx_solver *s, double *seconds, double *ops) {
    if (!s || !s->timers.on || !seconds || !ops) return OMX_ERR_ARG;
    double sum = 0.0;
    for (int i = 0; i < OMX_TIMER_ITEMS; ++i) {
        seconds[i] = s->timers.seconds[i];
        ops[i] = s->timers.ops[i];
        if (i < kT_TOTAL) sum += seconds[i];
    }
    // as Timer::show (timer.hpp:166-172): MISC = TOTAL - the other items
    seconds[OMX_TIMER_ITEMS - 1] = seconds[kT_TOTAL] - sum;
    return OMX_OK;
}

const char *omx_timer_name(int item) {
    return (item >= 0 && item < OMX_TIMER_ITEMS) ? kTimerNames[item] : nullptr;
}

}  // extern "C"
