// wv_internal.hpp -- kernel argument blocks and launchers shared between the
// HIP translation units and the C-ABI layer (wv_capi.hip).  Not installed.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/weavess_hip.h"

namespace wv {

// ------------------------------------------------------------ greedy search
struct SearchArgs {
    const float *base;        // N x dp, rows zero-padded to dp = roundup(d, 4)
    const float *queries;     // Q x dp, same padding
    const uint32_t *graph;    // N x deg_max
    const uint32_t *deg;      // N
    const uint32_t *entries;  // Q x L or null (seeded GenRandom)
    const uint32_t *L_per_q;  // Q or null
    const uint32_t *k_per_q;  // Q or null
    uint32_t *out_ids;        // Q x k_stride
    float *out_dist;          // Q x k_stride or null
    uint32_t *evals;          // Q or null
    uint32_t *hops;           // Q or null
    uint64_t seed;
    uint32_t N, dp, nch, deg_max;
    uint32_t Q, L, k, k_stride;
    uint32_t P;               // pool slots: every L rounded up to a multiple of 64
    uint32_t H;               // visited-table slots (power of two)
    uint32_t B;               // merge-buffer keys per wavefront (0: none; <= 256)
    uint32_t wave_lds;        // LDS bytes per wavefront
    const wv_term_model *model;  // learned termination (device copy) or null
    float *trace;             // Q x trace_cap x 8, or null
    uint32_t *trace_len;      // Q
    uint32_t *settle;         // Q
    uint32_t trace_cap;
    uint64_t *full;           // Q x C: the C smallest distinct keys evaluated (NSG candidates), or null
    uint32_t C;
    uint32_t *queue;          // persistent schedule: the work-queue counter (zeroed by the launch), or null
    const uint32_t *order;    // persistent schedule: query of each slot (Q), or null (slot i = query i)
};

// LDS bytes per wavefront; C > 0 adds the collected candidates and a block.
size_t search_wave_lds(uint32_t P, uint32_t H, uint32_t B, uint32_t C = 0);
// LDS bytes of one query served by W cooperating wavefronts.
size_t search_coop_lds(uint32_t P, uint32_t H, uint32_t B, uint32_t W);
// W (2 or 4) wavefronts of one workgroup per query (plain search only: no
// model, traces, collection or work queue).
hipError_t launch_greedy_search_coop(const SearchArgs &a, int chunks_per_lane, int waves_per_query, hipStream_t s);
// One wavefront per query (grid of ceil(Q / waves_per_block) workgroups),
// or, with a.queue, persistent wavefronts taking queries from a device work
// queue (about cu_count x the workgroups a CU holds).
hipError_t launch_greedy_search(const SearchArgs &a, int chunks_per_lane, int row_groups,
                                int waves_per_block, hipStream_t s, uint32_t cu_count = 256);

// ------------------------------------------------------------ MLP (wv_mlp.hip)
hipError_t mlp_grad_launch(const float *params, const float *mean, const float *inv_std, const float *X,
                           const float *y, const uint32_t *idx, uint32_t n_batch, float pos_weight,
                           float *partial, hipStream_t s);
hipError_t mlp_update_launch(const float *partial, uint32_t nblocks, float inv_n, float *params, float *m,
                             float *v, float *grad_out, float *loss_out, int adam, float lr, float beta1,
                             float beta2, float eps, float bc1, float bc2, hipStream_t s);
hipError_t mlp_predict_launch(const wv_term_model *m, const float *X, uint32_t n, float *out, hipStream_t s);
constexpr int MLP_PARTIAL_STRIDE = 324;

// ------------------------------------------------------------ exact kNN
struct ExactArgs {
    const float *base;     // N x d (unpadded rows of stride ld_base)
    const float *queries;  // Q x d (stride ld_query)
    uint32_t N, Q, d, ld_base, ld_query;
    uint32_t k;
    int64_t self_offset;   // >= 0: query q is base row q+self_offset, skipped
    uint32_t *out_ids;     // Q x k
    double *out_dist;      // Q x k
};
hipError_t launch_exact_knn(const ExactArgs &a, hipStream_t s);

// ------------------------------------------------------------ gather rate
// `blocks` one-wavefront workgroups, each scoring `batches` x 64 random rows
// of the N x dp table (dp <= 256) with the search's 64-row gather tile;
// `lds` bytes of dynamic LDS a workgroup (>= 768) set the residency.
hipError_t launch_gather_rate(const float *base, uint32_t N, uint32_t dp, uint32_t blocks, uint32_t batches,
                              uint32_t lds, uint64_t *out, hipStream_t s);
hipError_t launch_line_rate(const void *a, uint64_t bytes, uint32_t blocks, uint32_t iters, uint32_t *out,
                            hipStream_t s);

}  // namespace wv
