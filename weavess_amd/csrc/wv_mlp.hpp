// wv_mlp.hpp -- the learned-termination network on the device: search-state
// features and the BF16 inference step fused into the greedy search.
#pragma once

#include "wv_device.hpp"
#include "../../include/weavess_hip.h"

namespace wv {

constexpr int MLP_IN = 8, MLP_HID = 32, MLP_PARAMS = MLP_HID * MLP_IN + MLP_HID + MLP_HID + 1;  // 321
// The stop rule is evaluated after every MLP_STRIDE-th expansion (the
// features' window: "k-th distance 4 expansions earlier" is then the last
// evaluation's k-th distance).
constexpr uint32_t MLP_STRIDE = 4;

__device__ inline float bf16_round(float x) { return (float)(__bf16)x; }

// Feature vector g of a search state (see wv_term_model in weavess_hip.h):
// hops = expansions so far, kth/best = k-th/best pool distance, entry = best
// entry distance, kth4 = k-th distance 4 expansions earlier (the entries'
// k-th before the 4th), stale = expansions since a candidate last entered
// the top-k.  Every lane computes the same values.
__device__ inline void term_features(float hops, float kth, float best, float entry, float kth4,
                                     float stale, float L, float k, float f[MLP_IN]) {
    const float e = entry > 1e-30f ? entry : 1e-30f;
    f[0] = log2f(1.0f + hops);
    f[1] = kth / e;
    f[2] = best / e;
    f[3] = log2f(e);
    f[4] = (kth4 - kth) / e;
    f[5] = log2f(L);
    f[6] = log2f(k);
    f[7] = log2f(1.0f + stale);
}

// Logit of the network for features f, in BF16 products with FP32
// accumulation: lane j < 32 forms hidden unit j as
//   b1[j] + sum_i bf16(w1[j][i]) * bf16(x_i)   (i ascending),
// ReLU, times bf16(w2[j]) after rounding the activation to BF16; the 32
// products are summed by an xor butterfly (16, 8, 4, 2, 1) and b2 is added
// last.  x_i = (f_i - mean_i) * inv_std_i.  Returns the same value on every
// lane of the wavefront.
__device__ inline float term_logit(const wv_term_model *m, const float f[MLP_IN]) {
    const int lane = lane_id();
    float part = 0.0f;
    if (lane < MLP_HID) {
        float acc = m->b1[lane];
#pragma unroll
        for (int i = 0; i < MLP_IN; ++i) {
            const float x = bf16_round((f[i] - m->mean[i]) * m->inv_std[i]);
            acc = acc + bf16_round(m->w1[lane][i]) * x;
        }
        const float h = acc > 0.0f ? acc : 0.0f;
        part = bf16_round(m->w2[lane]) * bf16_round(h);
    }
    part = part + __shfl_xor(part, 16);
    part = part + __shfl_xor(part, 8);
    part = part + __shfl_xor(part, 4);
    part = part + __shfl_xor(part, 2);
    part = part + __shfl_xor(part, 1);
    return __shfl(part, 0) + m->b2;
}

// The network kept in registers for the fused search step: lane j < 32
// holds hidden unit j's weights (w1 and w2 rounded to BF16 once), every lane
// the standardisation constants.  term_logit_regs computes exactly
// term_logit's sums in the same order.
struct TermRegs {
    float w1[MLP_IN];
    float b1, w2;
    float mean[MLP_IN], inv_std[MLP_IN];
    float b2, threshold;
    uint32_t min_hops;
};

__device__ inline void term_load(const wv_term_model *m, TermRegs &r) {
    const int lane = lane_id();
    const int j = lane < MLP_HID ? lane : 0;
#pragma unroll
    for (int i = 0; i < MLP_IN; ++i) {
        r.w1[i] = bf16_round(m->w1[j][i]);
        r.mean[i] = m->mean[i];
        r.inv_std[i] = m->inv_std[i];
    }
    r.b1 = m->b1[j];
    r.w2 = bf16_round(m->w2[j]);
    r.b2 = m->b2;
    r.threshold = m->logit_threshold;
    r.min_hops = m->min_hops;
}

__device__ inline float term_logit_regs(const TermRegs &r, const float f[MLP_IN]) {
    const int lane = lane_id();
    float part = 0.0f;
    if (lane < MLP_HID) {
        float acc = r.b1;
#pragma unroll
        for (int i = 0; i < MLP_IN; ++i) {
            const float x = bf16_round((f[i] - r.mean[i]) * r.inv_std[i]);
            acc = acc + r.w1[i] * x;
        }
        const float h = acc > 0.0f ? acc : 0.0f;
        part = r.w2 * bf16_round(h);
    }
    part = part + __shfl_xor(part, 16);
    part = part + __shfl_xor(part, 8);
    part = part + __shfl_xor(part, 4);
    part = part + __shfl_xor(part, 2);
    part = part + __shfl_xor(part, 1);
    return __shfl(part, 0) + r.b2;
}

}  // namespace wv
