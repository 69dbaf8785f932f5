// omx_mlp.hip -- the learned per-step precision scheduler: an MLP
// 16-64-64-3 (FP32, ReLU) with forward, backward and Adam kernels, and the
// feature kernel that turns a panel step's statistics into its 16 inputs.
// No reference counterpart (the reference always uses FP16).
//
// Parameters are one flat FP32 array (OMX_MLP_NPARAMS):
//   W1[16][64] | b1[64] | W2[64][64] | b2[64] | W3[64][3] | b3[3]
// with W[i][j] the weight from input i to unit j.  Outputs z0, z1 are the
// logits of FP16 / FP8 for the step's Schur update; z2 is the FP8 tile-scale
// headroom in bits (rounded and clamped to 0..3).
//
// Contraction is off in this file: each sum is the same sequence of FP32
// roundings (bias first, inputs in order) as the CPU oracle, so the forward
// pass is bit-identical to it.
#include <cmath>

#include "omx_common.hpp"

#pragma clang fp contract(off)

namespace {

constexpr int kIn = 16, kH = 64, kOut = 3;
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kIn * kH;
constexpr int kW2 = kB1 + kH;
constexpr int kB2 = kW2 + kH * kH;
constexpr int kW3 = kB2 + kH;
constexpr int kB3 = kW3 + kH * kOut;
constexpr int kNP = kB3 + kOut;
static_assert(kNP == OMX_MLP_NPARAMS, "parameter count");
constexpr int kD = kOut + 2 * kH;  // per-sample backward deltas: dz, d2, d1

__device__ __forceinline__ void load_params(float *p, const float *__restrict__ params) {
    for (int i = threadIdx.x; i < kNP; i += blockDim.x) p[i] = params[i];
    __syncthreads();
}

__global__ void __launch_bounds__(256)
k_mlp_forward(int64_t batch, const float *__restrict__ params, const float *__restrict__ X, float *__restrict__ H1,
              float *__restrict__ H2, float *__restrict__ Z) {
    __shared__ float p[kNP];
    load_params(p, params);
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= batch) return;
    float x[kIn], h1[kH];
#pragma unroll
    for (int i = 0; i < kIn; ++i) x[i] = X[s * kIn + i];
    for (int j = 0; j < kH; ++j) {
        float a = p[kB1 + j];
#pragma unroll
        for (int i = 0; i < kIn; ++i) a += x[i] * p[kW1 + i * kH + j];
        h1[j] = a > 0.0f ? a : 0.0f;
        H1[s * kH + j] = h1[j];
    }
    float z[kOut];
#pragma unroll
    for (int j = 0; j < kOut; ++j) z[j] = p[kB3 + j];
    // h2 one unit at a time; its contribution to z is added after all of h2
    // is known, in the oracle's order (z_j = b3_j + sum_i h2_i W3[i][j])
    float h2[kH];
    for (int j = 0; j < kH; ++j) {
        float a = p[kB2 + j];
        for (int i = 0; i < kH; ++i) a += h1[i] * p[kW2 + i * kH + j];
        h2[j] = a > 0.0f ? a : 0.0f;
        H2[s * kH + j] = h2[j];
    }
#pragma unroll
    for (int j = 0; j < kOut; ++j) {
        float a = z[j];
        for (int i = 0; i < kH; ++i) a += h2[i] * p[kW3 + i * kOut + j];
        Z[s * kOut + j] = a;
    }
}

// Per-sample deltas of the mean loss: dz (3), d2 = dL/d(pre-activation 2)
// (64), d1 (64), and the sample's loss (FP64).
__global__ void __launch_bounds__(256)
k_mlp_delta(int64_t batch, const float *__restrict__ params, const float *__restrict__ H1,
            const float *__restrict__ H2, const float *__restrict__ Z, const int32_t *__restrict__ y,
            const float *__restrict__ t, const float *__restrict__ m, float *__restrict__ D,
            double *__restrict__ lossv) {
    __shared__ float p[kNP];
    load_params(p, params);
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= batch) return;
    const float inv = 1.0f / (float)batch;
    const float z0 = Z[s * kOut], z1 = Z[s * kOut + 1], z2 = Z[s * kOut + 2];
    const float mx = z0 > z1 ? z0 : z1;
    const float e0 = expf(z0 - mx), e1 = expf(z1 - mx);
    const float lse = mx + logf(e0 + e1);
    const float p0 = e0 / (e0 + e1), p1 = e1 / (e0 + e1);
    const int ys = y[s];
    const float diff = z2 - t[s];
    lossv[s] = (double)(lse - (ys ? z1 : z0)) + 0.5 * (double)(m[s] * diff * diff);
    float dz[kOut];
    dz[0] = (p0 - (ys ? 0.0f : 1.0f)) * inv;
    dz[1] = (p1 - (ys ? 1.0f : 0.0f)) * inv;
    dz[2] = m[s] * diff * inv;
    float *d = D + s * kD;
#pragma unroll
    for (int j = 0; j < kOut; ++j) d[j] = dz[j];
    float d2[kH];
    for (int i = 0; i < kH; ++i) {
        float a = 0.0f;
#pragma unroll
        for (int j = 0; j < kOut; ++j) a += p[kW3 + i * kOut + j] * dz[j];
        d2[i] = H2[s * kH + i] > 0.0f ? a : 0.0f;
        d[kOut + i] = d2[i];
    }
    for (int i = 0; i < kH; ++i) {
        float a = 0.0f;
        for (int j = 0; j < kH; ++j) a += p[kW2 + i * kH + j] * d2[j];
        d[kOut + kH + i] = H1[s * kH + i] > 0.0f ? a : 0.0f;
    }
}

// One thread per parameter: the sum over samples (in order) of its
// per-sample gradient.
__global__ void __launch_bounds__(256)
k_mlp_grad(int64_t batch, const float *__restrict__ X, const float *__restrict__ H1, const float *__restrict__ H2,
           const float *__restrict__ D, float *__restrict__ grad) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= kNP) return;
    float g = 0.0f;
    if (q < kB1) {
        const int i = q / kH, j = q % kH;
        for (int64_t s = 0; s < batch; ++s) g += X[s * kIn + i] * D[s * kD + kOut + kH + j];
    } else if (q < kW2) {
        const int j = q - kB1;
        for (int64_t s = 0; s < batch; ++s) g += D[s * kD + kOut + kH + j];
    } else if (q < kB2) {
        const int i = (q - kW2) / kH, j = (q - kW2) % kH;
        for (int64_t s = 0; s < batch; ++s) g += H1[s * kH + i] * D[s * kD + kOut + j];
    } else if (q < kW3) {
        const int j = q - kB2;
        for (int64_t s = 0; s < batch; ++s) g += D[s * kD + kOut + j];
    } else if (q < kB3) {
        const int i = (q - kW3) / kOut, j = (q - kW3) % kOut;
        for (int64_t s = 0; s < batch; ++s) g += H2[s * kH + i] * D[s * kD + j];
    } else {
        const int j = q - kB3;
        for (int64_t s = 0; s < batch; ++s) g += D[s * kD + j];
    }
    grad[q] = g;
}

__global__ void k_loss_reduce(int64_t batch, const double *__restrict__ lossv, float *__restrict__ loss) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = 0.0;
    for (int64_t s = 0; s < batch; ++s) a += lossv[s];
    *loss = (float)(a / (double)batch);
}

__global__ void k_mlp_adam(int64_t np, float *__restrict__ params, const float *__restrict__ grad,
                           float *__restrict__ mom, float *__restrict__ vel, float lr, float b1, float b2, float eps,
                           float c1, float c2) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const float g = grad[q];
    const float mq = b1 * mom[q] + (1.0f - b1) * g;
    const float vq = b2 * vel[q] + (1.0f - b2) * g * g;
    mom[q] = mq;
    vel[q] = vq;
    params[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

__global__ void k_mlp_sgd(int64_t np, float *__restrict__ params, const float *__restrict__ grad,
                          float *__restrict__ vel, float lr, float mu) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    float g = grad[q];
    if (vel) {
        g = mu * vel[q] + g;
        vel[q] = g;
    }
    params[q] -= lr * g;
}

// Statistics of log2 of the non-zero tile maxima of a panel: max, min,
// mean, standard deviation (thread-strided partial sums, then a fixed-order
// tree in LDS).
struct LogStats {
    float mx, mn, mean, sd;
};

__device__ LogStats log_stats(int64_t cnt, const float *__restrict__ v, float *red) {
    const int t = threadIdx.x;
    float mx = -1e30f, mn = 1e30f, s1 = 0.0f, s2 = 0.0f, c = 0.0f;
    for (int64_t i = t; i < cnt; i += blockDim.x) {
        const float a = v[i];
        if (a > 0.0f && a < __builtin_inff()) {
            const float l = log2f(a);
            mx = fmaxf(mx, l);
            mn = fminf(mn, l);
            s1 += l;
            s2 += l * l;
            c += 1.0f;
        }
    }
    red[t] = mx;
    red[256 + t] = mn;
    red[512 + t] = s1;
    red[768 + t] = s2;
    red[1024 + t] = c;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            red[t] = fmaxf(red[t], red[t + w]);
            red[256 + t] = fminf(red[256 + t], red[256 + t + w]);
            red[512 + t] += red[512 + t + w];
            red[768 + t] += red[768 + t + w];
            red[1024 + t] += red[1024 + t + w];
        }
        __syncthreads();
    }
    LogStats r;
    const float n = red[1024];
    if (n > 0.0f) {
        r.mx = red[0];
        r.mn = red[256];
        r.mean = red[512] / n;
        r.sd = sqrtf(fmaxf(red[768] / n - r.mean * r.mean, 0.0f));
    } else {
        r.mx = r.mn = r.mean = -64.0f;
        r.sd = 0.0f;
    }
    __syncthreads();
    return r;
}

__device__ __forceinline__ float clip8(float v) {
    if (!(v == v)) return 0.0f;
    return fminf(fmaxf(v, -8.0f), 8.0f);
}

// The 16 inputs of one panel step (one workgroup of 256 threads).
__global__ void __launch_bounds__(256)
k_sched_features(int64_t nl, const float *__restrict__ lmax, int64_t nr, const float *__restrict__ rmax,
                 const float *__restrict__ Akk, int64_t lda, int64_t b, int64_t step, int64_t nb, int64_t n,
                 int prev, float *__restrict__ out) {
    __shared__ float red[5 * 256];
    const LogStats L = log_stats(nl, lmax, red);
    const LogStats R = log_stats(nr, rmax, red);
    // |u_ii| of the factored diagonal block
    const int t = threadIdx.x;
    float mx = -1e30f, mn = 1e30f;
    for (int64_t i = t; i < b; i += blockDim.x) {
        const float u = fabsf(Akk[i + i * lda]);
        const float l = (u > 0.0f && u < __builtin_inff()) ? log2f(u) : -64.0f;
        mx = fmaxf(mx, l);
        mn = fminf(mn, l);
    }
    red[t] = mx;
    red[256 + t] = mn;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            red[t] = fmaxf(red[t], red[t + w]);
            red[256 + t] = fminf(red[256 + t], red[256 + t + w]);
        }
        __syncthreads();
    }
    if (t != 0) return;
    const float umx = red[0], umn = red[256];
    const int64_t rest = n - (step + 1) * b;
    float f[16];
    f[0] = L.mx / 16.0f;
    f[1] = L.mn / 16.0f;
    f[2] = L.mean / 16.0f;
    f[3] = L.sd / 8.0f;
    f[4] = R.mx / 16.0f;
    f[5] = R.mn / 16.0f;
    f[6] = R.mean / 16.0f;
    f[7] = R.sd / 8.0f;
    f[8] = umx / 16.0f;
    f[9] = umn / 16.0f;
    f[10] = (umx - umn) / 16.0f;
    f[11] = (float)step / (float)nb;
    f[12] = log2f((float)(rest > 0 ? rest : 1) / (float)n) / 8.0f;
    f[13] = (L.mx - umx) / 16.0f;
    f[14] = (R.mx - umx) / 16.0f;
    f[15] = (float)prev;
#pragma unroll
    for (int i = 0; i < 16; ++i) out[i] = clip8(f[i]);
}

}  // namespace

extern "C" {

int omx_mlp_forward_f32(int64_t batch, const float *params, const float *X, float *H1, float *H2, float *Z,
                        void *stream) {
    if (batch <= 0 || !params || !X || !H1 || !H2 || !Z) return OMX_ERR_ARG;
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)omx_cdiv(batch, 256)), dim3(256), 0, omx_stream(stream),
                       batch, params, X, H1, H2, Z);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

size_t omx_mlp_backward_workspace_bytes(int64_t batch) {
    if (batch <= 0) return 0;
    return (size_t)batch * (kD * sizeof(float) + sizeof(double));
}

int omx_mlp_backward_f32(int64_t batch, const float *params, const float *X, const float *H1, const float *H2,
                         const float *Z, const int32_t *y, const float *t, const float *m, float *grad, float *loss,
                         void *work, size_t work_bytes, void *stream) {
    if (batch <= 0 || !params || !X || !H1 || !H2 || !Z || !y || !t || !m || !grad || !loss) return OMX_ERR_ARG;
    if (!work || work_bytes < omx_mlp_backward_workspace_bytes(batch)) return OMX_ERR_WORKSPACE;
    hipStream_t st = omx_stream(stream);
    double *lossv = static_cast<double *>(work);
    float *D = reinterpret_cast<float *>(lossv + batch);
    hipLaunchKernelGGL(k_mlp_delta, dim3((unsigned)omx_cdiv(batch, 256)), dim3(256), 0, st, batch, params, H1, H2,
                       Z, y, t, m, D, lossv);
    OMX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)omx_cdiv(kNP, 256)), dim3(256), 0, st, batch, X, H1, H2, D, grad);
    OMX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_loss_reduce, dim3(1), dim3(64), 0, st, batch, lossv, loss);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

int omx_mlp_adam_f32(int64_t nparams, float *params, const float *grad, float *mom, float *vel, float lr,
                     float beta1, float beta2, float eps, int step, void *stream) {
    if (nparams <= 0 || !params || !grad || !mom || !vel || step < 1 || !(lr > 0) || !(beta1 >= 0) ||
        !(beta1 < 1) || !(beta2 >= 0) || !(beta2 < 1) || !(eps > 0))
        return OMX_ERR_ARG;
    const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_mlp_adam, dim3((unsigned)omx_cdiv(nparams, 256)), dim3(256), 0, omx_stream(stream),
                       nparams, params, grad, mom, vel, lr, beta1, beta2, eps, c1, c2);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

int omx_mlp_sgd_f32(int64_t nparams, float *params, const float *grad, float *vel, float lr, float momentum,
                    void *stream) {
    if (nparams <= 0 || !params || !grad || !(lr > 0) || !(momentum >= 0) || !(momentum < 1) ||
        (momentum > 0 && !vel))
        return OMX_ERR_ARG;
    hipLaunchKernelGGL(k_mlp_sgd, dim3((unsigned)omx_cdiv(nparams, 256)), dim3(256), 0, omx_stream(stream),
                       nparams, params, grad, momentum > 0 ? vel : nullptr, lr, momentum);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

int omx_sched_features_f32(int64_t ntiles_l, const float *lmax, int64_t ntiles_r, const float *rmax,
                           const float *Akk, int64_t lda, int64_t b, int64_t step, int64_t nb, int64_t n, int prev,
                           float *features, void *stream) {
    if (ntiles_l < 0 || ntiles_r < 0 || (ntiles_l > 0 && !lmax) || (ntiles_r > 0 && !rmax) || !Akk || b <= 0 ||
        lda < b || step < 0 || nb <= 0 || n <= 0 || !features)
        return OMX_ERR_ARG;
    hipLaunchKernelGGL(k_sched_features, dim3(1), dim3(256), 0, omx_stream(stream), ntiles_l, lmax, ntiles_r, rmax,
                       Akk, lda, b, step, nb, n, prev, features);
    OMX_LAUNCH_CHECK();
    return OMX_OK;
}

}  // extern "C"
