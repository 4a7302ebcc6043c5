// SPDX-License-Identifier: MIT
//
// The learned scheduler's kernels: per-world features from the park's state and the path classes its last tick
// recorded, and a 12 -> 64 -> 64 -> 8 FP32 multilayer perceptron (ReLU hidden layers, softmax cross-entropy) with its
// forward pass, backward pass and Adam step. It predicts each world's cost bucket for the next tick; the park then
// regroups worlds of like cost into the same workgroups. Nothing here touches the tick's results: a world's bits do not
// depend on which workgroup runs it.
//
// Determinism: no atomics; every reduction over samples is one thread summing in sample order, so a run repeats bit
// for bit; exp, log and pow are the glibc transcriptions the tick uses (xb_maths_*.h), the same on the host and the
// GPU. The network is not part of the tick's exactness contract: its multiply-adds are explicit fused steps.
#include <hip/hip_runtime.h>

#define XB_HD __device__
#define XB_TABLE __constant__
#include "xb_modes_tick.h"
#include "xb_maths_float.h"
#include "../../include/xb.h"

namespace {

constexpr int NI = XB_MLP_IN, NH = XB_MLP_HIDDEN, NO = XB_MLP_OUT;
constexpr int OFF_W1 = 0, OFF_B1 = NI * NH, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH, OFF_W3 = OFF_B2 + NH,
              OFF_B3 = OFF_W3 + NH * NO;
static_assert(OFF_B3 + NO == XB_MLP_PARAMS, "parameter layout");

// the integrator-only weight of each path class, from the double functions and divisions each calls
__constant__ float kPathCost[8] = {0.0f, 1.0f, 2.0f, 3.0f, 8.0f, 9.0f, 7.0f, 8.0f};

__global__ void xb_features_kernel(xb_modes_state s, const int32_t* __restrict__ wstart,
                                   const int32_t* __restrict__ wcount, int32_t n_worlds, const uint8_t* path,
                                   float* __restrict__ feat) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n_worlds) return;
    XbPark P;
    P.f = s.f64, P.q = s.i64, P.b = s.u8, P.stride = s.stride;
    const int lo = wstart[w], cnt = wcount[w];
    float hist[5] = {0.f, 0.f, 0.f, 0.f, 0.f};       // stop+inert, goto, follow, orbit, warp
    float aligning = 0.f, idle = 0.f, cost = 0.f;
    unsigned seen = 0u;                               // the path classes the world's rows executed
    double cx = 0.0, cy = 0.0, cz = 0.0;
    for (int k = 0; k < cnt; ++k) {
        const int64_t r = lo + k;
        const int64_t m = XB_Q(P, XB_I_MODE, r);
        const bool live = XB_B(P, XB_B_FREE, r) && !XB_B(P, XB_B_MORIBUND, r);
        idle += live ? 0.f : 1.f;
        const int h = m == XB_MODE_GOTO ? 1 : m == XB_MODE_FOLLOW ? 2 : m == XB_MODE_ORBIT ? 3 : m == XB_MODE_WARP ? 4 : 0;
        hist[h] += 1.f;
        if (m == XB_MODE_WARP && XB_Q(P, XB_I_EFFECT_STAMP, r) < 0) aligning += 1.f;
        if (path) cost += kPathCost[path[r] & 7], seen |= 1u << (path[r] & 7);
        cx += XB_F(P, XB_F_POS, r), cy += XB_F(P, XB_F_POS + 1, r), cz += XB_F(P, XB_F_POS + 2, r);
    }
    const float inv = cnt > 0 ? 1.f / (float)cnt : 0.f;
    const double dinv = cnt > 0 ? 1.0 / (double)cnt : 0.0;
    cx *= dinv, cy *= dinv, cz *= dinv;
    double dmin = 1e300, dsum = 0.0;
    for (int k = 0; k < cnt; ++k) {
        const int64_t r = lo + k;
        const double dx = XB_F(P, XB_F_POS, r) - cx, dy = XB_F(P, XB_F_POS + 1, r) - cy, dz = XB_F(P, XB_F_POS + 2, r) - cz;
        const double d = __builtin_sqrt(dx * dx + dy * dy + dz * dz);
        dsum += d;
        dmin = d < dmin ? d : dmin;
    }
    float* o = feat + (size_t)w * NI;
    for (int h = 0; h < 5; ++h) o[h] = hist[h] * inv;
    o[5] = (float)cnt / 32.f;
    o[6] = aligning * inv;
    o[7] = idle * inv;
    o[8] = cnt > 0 ? (float)(xb_log(1.0 + dmin) / 23.025850929940457) : 0.f;
    o[9] = cnt > 0 ? (float)(xb_log(1.0 + dsum * dinv) / 23.025850929940457) : 0.f;
    float uni = 0.f;
    for (int c = 0; c < 8; ++c) uni += (seen >> c) & 1u ? kPathCost[c] : 0.f;
    o[10] = uni / 38.f;
    o[11] = cost * inv / 9.f;
}

// One sample a thread; the parameters staged in LDS once per workgroup.
__global__ void __launch_bounds__(256) xb_mlp_forward_kernel(const float* __restrict__ params,
                                                             const float* __restrict__ x, int32_t n,
                                                             float* __restrict__ h1, float* __restrict__ h2,
                                                             float* __restrict__ logits) {
    __shared__ float p[XB_MLP_PARAMS];
    for (int i = threadIdx.x; i < XB_MLP_PARAMS; i += blockDim.x) p[i] = params[i];
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float in[NI], a1[NH], a2[NH];
    for (int i = 0; i < NI; ++i) in[i] = x[(size_t)s * NI + i];
    for (int j = 0; j < NH; ++j) {
        float acc = p[OFF_B1 + j];
        for (int i = 0; i < NI; ++i) acc = __builtin_fmaf(in[i], p[OFF_W1 + i * NH + j], acc);
        a1[j] = acc > 0.f ? acc : 0.f;
    }
    for (int j = 0; j < NH; ++j) {
        float acc = p[OFF_B2 + j];
        for (int i = 0; i < NH; ++i) acc = __builtin_fmaf(a1[i], p[OFF_W2 + i * NH + j], acc);
        a2[j] = acc > 0.f ? acc : 0.f;
    }
    for (int k = 0; k < NO; ++k) {
        float acc = p[OFF_B3 + k];
        for (int i = 0; i < NH; ++i) acc = __builtin_fmaf(a2[i], p[OFF_W3 + i * NO + k], acc);
        logits[(size_t)s * NO + k] = acc;
    }
    if (h1)
        for (int j = 0; j < NH; ++j) h1[(size_t)s * NH + j] = a1[j];
    if (h2)
        for (int j = 0; j < NH; ++j) h2[(size_t)s * NH + j] = a2[j];
}

// Per sample: the loss's gradient with respect to the logits (softmax minus one-hot, over n), and back through the
// hidden layers (ReLU masks from the saved activations).
__global__ void __launch_bounds__(256) xb_mlp_delta_kernel(const float* __restrict__ params,
                                                           const float* __restrict__ h1, const float* __restrict__ h2,
                                                           const float* __restrict__ logits,
                                                           const int32_t* __restrict__ labels, int32_t n,
                                                           float* __restrict__ dl, float* __restrict__ d2,
                                                           float* __restrict__ d1, float* __restrict__ loss) {
    __shared__ float p[XB_MLP_PARAMS];
    for (int i = threadIdx.x; i < XB_MLP_PARAMS; i += blockDim.x) p[i] = params[i];
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float* z = logits + (size_t)s * NO;
    float mx = z[0];
    for (int k = 1; k < NO; ++k) mx = z[k] > mx ? z[k] : mx;
    float e[NO], sum = 0.f;
    for (int k = 0; k < NO; ++k) {
        e[k] = xb_expf(z[k] - mx);
        sum += e[k];
    }
    const int y = (uint32_t)labels[s] < (uint32_t)NO ? labels[s] : 0;   // the host checks labels; kept in bounds
    const float invn = 1.f / (float)n;
    float g3[NO];
    for (int k = 0; k < NO; ++k) {
        g3[k] = (e[k] / sum - (k == y ? 1.f : 0.f)) * invn;
        dl[(size_t)s * NO + k] = g3[k];
    }
    loss[s] = (xb_logf(sum) + mx - z[y]) * invn;
    float g2[NH];
    for (int i = 0; i < NH; ++i) {
        float acc = 0.f;
        for (int k = 0; k < NO; ++k) acc = __builtin_fmaf(g3[k], p[OFF_W3 + i * NO + k], acc);
        g2[i] = h2[(size_t)s * NH + i] > 0.f ? acc : 0.f;
        d2[(size_t)s * NH + i] = g2[i];
    }
    for (int i = 0; i < NH; ++i) {
        float acc = 0.f;
        for (int j = 0; j < NH; ++j) acc = __builtin_fmaf(g2[j], p[OFF_W2 + i * NH + j], acc);
        d1[(size_t)s * NH + i] = h1[(size_t)s * NH + i] > 0.f ? acc : 0.f;
    }
}

// One parameter a thread, one run of samples a grid row (blockIdx.y): the parameter's gradient summed over the run in
// sample order into work[run][param].
__global__ void xb_mlp_grad_kernel(const float* __restrict__ x, const float* __restrict__ h1,
                                   const float* __restrict__ h2, const float* __restrict__ dl,
                                   const float* __restrict__ d2, const float* __restrict__ d1, int32_t n,
                                   float* __restrict__ work) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= XB_MLP_PARAMS) return;
    const int c = blockIdx.y;
    const int per = (n + XB_MLP_CHUNKS - 1) / XB_MLP_CHUNKS;
    const int s0 = c * per, s1 = s0 + per < n ? s0 + per : n;
    const float *a = nullptr, *d = nullptr;
    int na = 0, nd = 0, i = 0, j = 0;
    if (q < OFF_B1) {
        a = x, na = NI, d = d1, nd = NH, i = q / NH, j = q % NH;
    } else if (q < OFF_W2) {
        d = d1, nd = NH, j = q - OFF_B1;
    } else if (q < OFF_B2) {
        a = h1, na = NH, d = d2, nd = NH, i = (q - OFF_W2) / NH, j = (q - OFF_W2) % NH;
    } else if (q < OFF_W3) {
        d = d2, nd = NH, j = q - OFF_B2;
    } else if (q < OFF_B3) {
        a = h2, na = NH, d = dl, nd = NO, i = (q - OFF_W3) / NO, j = (q - OFF_W3) % NO;
    } else {
        d = dl, nd = NO, j = q - OFF_B3;
    }
    float acc = 0.f;
    for (int s = s0; s < s1; ++s) {
        const float g = d[(size_t)s * nd + j];
        acc = a ? __builtin_fmaf(a[(size_t)s * na + i], g, acc) : acc + g;
    }
    work[(size_t)c * XB_MLP_PARAMS + q] = acc;
}

// The runs' partial gradients summed in run order.
__global__ void xb_mlp_grad_sum_kernel(const float* __restrict__ work, float* __restrict__ grads) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= XB_MLP_PARAMS) return;
    float acc = 0.f;
    for (int c = 0; c < XB_MLP_CHUNKS; ++c) acc += work[(size_t)c * XB_MLP_PARAMS + q];
    grads[q] = acc;
}

__global__ void xb_mlp_adam_kernel(float* __restrict__ params, const float* __restrict__ grads, float* __restrict__ m,
                                   float* __restrict__ v, int32_t step, float lr) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= XB_MLP_PARAMS) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float g = grads[q];
    m[q] = b1 * m[q] + (1.f - b1) * g;
    v[q] = b2 * v[q] + (1.f - b2) * g * g;
    const float mh = m[q] / (float)(1.0 - xb_pow((double)b1, (double)step));
    const float vh = v[q] / (float)(1.0 - xb_pow((double)b2, (double)step));
    params[q] -= lr * mh / (__builtin_sqrtf(vh) + eps);
}

int rc_of() {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : (int)e;
}

}  // namespace

extern "C" {

int xb_world_features(const xb_modes_state* state, const int32_t* world_start, const int32_t* world_count,
                      int32_t n_worlds, const uint8_t* path, float* features, void* stream) {
    if (!state || n_worlds < 0 || !world_start || !world_count || !features) return -1;
    if (n_worlds == 0) return 0;
    xb_features_kernel<<<(n_worlds + 255) / 256, 256, 0, (hipStream_t)stream>>>(*state, world_start, world_count,
                                                                                 n_worlds, path, features);
    return rc_of();
}

int xb_mlp_forward(const float* params, const float* x, int32_t n, float* h1, float* h2, float* logits, void* stream) {
    if (!params || !x || !logits || n < 0) return -1;
    if (n == 0) return 0;
    xb_mlp_forward_kernel<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(params, x, n, h1, h2, logits);
    return rc_of();
}

int xb_mlp_backward(const float* params, const float* x, const float* h1, const float* h2, const float* logits,
                    const int32_t* labels, int32_t n, float* dl, float* d2, float* d1, float* loss, float* work,
                    float* grads, void* stream) {
    if (!params || !x || !h1 || !h2 || !logits || !labels || !dl || !d2 || !d1 || !loss || !work || !grads || n <= 0)
        return -1;
    hipStream_t st = (hipStream_t)stream;
    xb_mlp_delta_kernel<<<(n + 255) / 256, 256, 0, st>>>(params, h1, h2, logits, labels, n, dl, d2, d1, loss);
    int rc = rc_of();
    if (rc) return rc;
    xb_mlp_grad_kernel<<<dim3((XB_MLP_PARAMS + 127) / 128, XB_MLP_CHUNKS), 128, 0, st>>>(x, h1, h2, dl, d2, d1, n,
                                                                                         work);
    rc = rc_of();
    if (rc) return rc;
    xb_mlp_grad_sum_kernel<<<(XB_MLP_PARAMS + 127) / 128, 128, 0, st>>>(work, grads);
    return rc_of();
}

int xb_mlp_adam(float* params, const float* grads, float* m, float* v, int32_t step, float lr, void* stream) {
    if (!params || !grads || !m || !v || step < 1) return -1;
    xb_mlp_adam_kernel<<<(XB_MLP_PARAMS + 127) / 128, 128, 0, (hipStream_t)stream>>>(params, grads, m, v, step, lr);
    return rc_of();
}

}  // extern "C"
