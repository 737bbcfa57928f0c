// Router network kernels: a 2-layer MLP (FP32) with ReLU hidden layer.
//
//   hidden = relu(W1 x + b1)       W1: H x F, b1: H
//   out    = W2 hidden + b2        W2: K x H, b2: K
//
// Parameters live in one flat array [W1, b1, W2, b2].  The backward kernel
// computes every parameter's gradient with one thread summing over the
// samples in index order, so gradients are deterministic (no atomics) and
// match a sequential FP32 sum; the router trains on mini-batches, where this
// is cheap.  The loss kernel forms the gradient of a masked mean squared
// error, so a whole training step (forward, loss gradient, backward, Adam)
// runs on the device.
#pragma once

#include "sable_device.h"

extern "C" __global__ void __launch_bounds__(256)
sb_mlp_forward_kernel(const float *__restrict__ P, int F, int H, int K, const float *__restrict__ X, int n,
                      float *__restrict__ out, float *__restrict__ hidden)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *W1 = P;
    const float *b1 = W1 + (int64_t)H * F;
    const float *W2 = b1 + H;
    const float *b2 = W2 + (int64_t)K * H;
    const float *x = X + (int64_t)i * F;
    float *h = hidden + (int64_t)i * H;
    for (int j = 0; j < H; ++j) {
        float s = b1[j];
        for (int f = 0; f < F; ++f) s += W1[j * F + f] * x[f];
        h[j] = s > 0.f ? s : 0.f;
    }
    for (int k = 0; k < K; ++k) {
        float s = b2[k];
        for (int j = 0; j < H; ++j) s += W2[k * H + j] * h[j];
        out[(int64_t)i * K + k] = s;
    }
}

extern "C" __global__ void __launch_bounds__(256)
sb_mlp_backward_kernel(const float *__restrict__ P, int F, int H, int K, const float *__restrict__ X, int n,
                       const float *__restrict__ hidden, const float *__restrict__ dout, float *__restrict__ G)
{
    const int64_t nW1 = (int64_t)H * F, nb1 = H, nW2 = (int64_t)K * H, nb2 = K;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nW1 + nb1 + nW2 + nb2) return;
    const float *W2 = P + nW1 + nb1;
    float g = 0.f;
    if (p < nW1 + nb1) {
        // first layer: dL/dW1[j][f] = sum_i dh[i][j] x[i][f],  dL/db1[j] = sum_i dh[i][j]
        const int j = p < nW1 ? (int)(p / F) : (int)(p - nW1);
        const int f = p < nW1 ? (int)(p % F) : -1;
        for (int i = 0; i < n; ++i) {
            const float hv = hidden[(int64_t)i * H + j];
            if (hv <= 0.f) continue;
            float d = 0.f;
            for (int k = 0; k < K; ++k) d += dout[(int64_t)i * K + k] * W2[k * H + j];
            g += f >= 0 ? d * X[(int64_t)i * F + f] : d;
        }
    } else if (p < nW1 + nb1 + nW2) {
        const int64_t q = p - nW1 - nb1;
        const int k = (int)(q / H), j = (int)(q % H);
        for (int i = 0; i < n; ++i) g += dout[(int64_t)i * K + k] * hidden[(int64_t)i * H + j];
    } else {
        const int k = (int)(p - nW1 - nb1 - nW2);
        for (int i = 0; i < n; ++i) g += dout[(int64_t)i * K + k];
    }
    G[p] = g;
}

// dout = scale * 2 * (out - target) * mask   (mask 0/1 per output)
extern "C" __global__ void __launch_bounds__(256)
sb_mlp_mse_grad_kernel(const float *__restrict__ out, const float *__restrict__ target,
                       const float *__restrict__ mask, int64_t n, float scale, float *__restrict__ dout)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dout[i] = scale * 2.f * (out[i] - target[i]) * mask[i];
}

extern "C" __global__ void __launch_bounds__(256)
sb_adam_kernel(float *__restrict__ P, const float *__restrict__ G, float *__restrict__ m, float *__restrict__ v,
               int64_t n, float lr, float b1, float b2, float eps, int step)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = G[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float mhat = mi / (1.f - powf(b1, (float)step));
    const float vhat = vi / (1.f - powf(b2, (float)step));
    P[i] -= lr * mhat / (sqrtf(vhat) + eps);
}
