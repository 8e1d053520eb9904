// wv_mlp.hip -- training and batch inference of the learned-termination MLP.
//
// Forward (FP32): x = (g - mean) * inv_std, z1 = W1 x + b1, h = relu(z1),
// z = w2 . h + b2.  Loss: weighted binary cross-entropy on the logit,
// w * (max(z,0) - z*y + log(1 + exp(-|z|))), w = pos_weight for y = 1.
// Backward: dz = w * (sigmoid(z) - y); dw2 = dz h; db2 = dz;
// dz1 = dz w2 [z1 > 0]; dW1 = dz1 x^T; db1 = dz1.
//
// One wavefront takes 64 samples (one per lane), parks x, dz1, h and dz in
// LDS, then lane p sums parameter p's gradient over the 64 samples in sample
// order; blocks' partial sums are added in block order by the update kernel,
// so gradients are deterministic.  The update is Adam.
#include "wv_device.hpp"
#include "wv_internal.hpp"
#include "wv_mlp.hpp"

namespace wv {
namespace {

constexpr int PSTRIDE = 324;  // floats per partial row: 321 gradients, loss, pad

__global__ void __launch_bounds__(64) mlp_grad_kernel(const float *params, const float *mean,
                                                      const float *inv_std, const float *X,
                                                      const float *y, const uint32_t *idx,
                                                      uint32_t n_batch, float pos_weight,
                                                      float *partial) {
    __shared__ float sx[64][MLP_IN];
    __shared__ float sdz1[64][MLP_HID + 1];
    __shared__ float sh[64][MLP_HID + 1];
    __shared__ float sdz[64];
    __shared__ float sloss[64];
    const int lane = threadIdx.x;
    const uint32_t s = blockIdx.x * 64 + lane;
    const bool live = s < n_batch;
    const float *W1 = params, *b1 = params + 256, *w2 = params + 288, b2 = params[320];
    float x[MLP_IN];
    const uint32_t row = live ? (idx ? idx[s] : s) : 0;
#pragma unroll
    for (int i = 0; i < MLP_IN; ++i) x[i] = live ? (X[(size_t)row * MLP_IN + i] - mean[i]) * inv_std[i] : 0.f;
    float z = b2;
    float h[MLP_HID], z1[MLP_HID];
#pragma unroll
    for (int j = 0; j < MLP_HID; ++j) {
        float acc = b1[j];
#pragma unroll
        for (int i = 0; i < MLP_IN; ++i) acc = acc + W1[j * MLP_IN + i] * x[i];
        z1[j] = acc;
        h[j] = acc > 0.f ? acc : 0.f;
    }
#pragma unroll
    for (int j = 0; j < MLP_HID; ++j) z = z + w2[j] * h[j];
    const float t = live ? y[row] : 0.f;
    const float w = live ? (t > 0.5f ? pos_weight : 1.0f) : 0.f;
    const float sig = 1.0f / (1.0f + expf(-z));
    const float dz = w * (sig - t);
    const float loss = w * (fmaxf(z, 0.f) - z * t + log1pf(expf(-fabsf(z))));
#pragma unroll
    for (int i = 0; i < MLP_IN; ++i) sx[lane][i] = x[i];
#pragma unroll
    for (int j = 0; j < MLP_HID; ++j) {
        sh[lane][j] = h[j];
        sdz1[lane][j] = z1[j] > 0.f ? dz * w2[j] : 0.f;
    }
    sdz[lane] = dz;
    sloss[lane] = loss;
    __syncthreads();
    float *out = partial + (size_t)blockIdx.x * PSTRIDE;
    for (int p = lane; p < MLP_PARAMS + 1; p += 64) {
        float acc = 0.f;
        if (p < 256) {
            const int j = p / MLP_IN, i = p % MLP_IN;
            for (int r = 0; r < 64; ++r) acc = acc + sdz1[r][j] * sx[r][i];
        } else if (p < 288) {
            for (int r = 0; r < 64; ++r) acc = acc + sdz1[r][p - 256];
        } else if (p < 320) {
            for (int r = 0; r < 64; ++r) acc = acc + sdz[r] * sh[r][p - 288];
        } else if (p == 320) {
            for (int r = 0; r < 64; ++r) acc = acc + sdz[r];
        } else {
            for (int r = 0; r < 64; ++r) acc = acc + sloss[r];
        }
        out[p] = acc;
    }
}

// Sum block partials in block order; mean over n; optional Adam step.
__global__ void __launch_bounds__(384) mlp_update_kernel(const float *partial, uint32_t nblocks, float inv_n,
                                                         float *params, float *m, float *v, float *grad_out,
                                                         float *loss_out, int adam, float lr, float beta1,
                                                         float beta2, float eps, float bc1, float bc2) {
    const int p = threadIdx.x;
    if (p > MLP_PARAMS) return;
    float acc = 0.f;
    for (uint32_t b = 0; b < nblocks; ++b) acc = acc + partial[(size_t)b * PSTRIDE + p];
    const float g = acc * inv_n;
    if (p == MLP_PARAMS) {
        if (loss_out) *loss_out = g;
        return;
    }
    if (grad_out) grad_out[p] = g;
    if (adam) {
        const float mm = beta1 * m[p] + (1.f - beta1) * g;
        const float vv = beta2 * v[p] + (1.f - beta2) * g * g;
        m[p] = mm;
        v[p] = vv;
        params[p] = params[p] - lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    }
}

__global__ void mlp_predict_kernel(const wv_term_model *m, const float *X, uint32_t n, float *out) {
    const uint32_t s = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (s >= n) return;
    float f[MLP_IN];
#pragma unroll
    for (int i = 0; i < MLP_IN; ++i) f[i] = X[(size_t)s * MLP_IN + i];
    const float z = term_logit(m, f);
    if (lane_id() == 0) out[s] = z;
}

}  // namespace

hipError_t mlp_grad_launch(const float *params, const float *mean, const float *inv_std, const float *X,
                           const float *y, const uint32_t *idx, uint32_t n_batch, float pos_weight,
                           float *partial, hipStream_t s) {
    const uint32_t blocks = (n_batch + 63) / 64;
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks), dim3(64), 0, s, params, mean, inv_std, X, y, idx,
                       n_batch, pos_weight, partial);
    return hipGetLastError();
}

hipError_t mlp_update_launch(const float *partial, uint32_t nblocks, float inv_n, float *params, float *m,
                             float *v, float *grad_out, float *loss_out, int adam, float lr, float beta1,
                             float beta2, float eps, float bc1, float bc2, hipStream_t s) {
    hipLaunchKernelGGL(mlp_update_kernel, dim3(1), dim3(384), 0, s, partial, nblocks, inv_n, params, m, v,
                       grad_out, loss_out, adam, lr, beta1, beta2, eps, bc1, bc2);
    return hipGetLastError();
}

hipError_t mlp_predict_launch(const wv_term_model *m, const float *X, uint32_t n, float *out, hipStream_t s) {
    const uint32_t blocks = (n + 3) / 4;
    hipLaunchKernelGGL(mlp_predict_kernel, dim3(blocks), dim3(256), 0, s, m, X, n, out);
    return hipGetLastError();
}

}  // namespace wv
