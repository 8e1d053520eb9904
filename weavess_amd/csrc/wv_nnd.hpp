// wv_nnd.hpp -- NN-Descent build on a device-resident index (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/weavess_hip.h"

namespace wv {

struct IndexView {
    const float *base;  // N x dp, zero padded
    uint32_t N, d, dp;
    hipStream_t stream;
};

// Implemented in wv_capi.hip.
IndexView index_view(wv_index *ix);
int index_alloc_graph(wv_index *ix, uint32_t deg_max);
int index_set_error(wv_index *ix, int code, const char *msg);

// Per-round graph recall of control nodes (the reference's debug monitor):
// after round r, hits[r * n_ctrl + i] = |pool(ctrl[i]) ∩ truth row i|.
struct NndMonitor {
    const uint32_t *ctrl;   // n_ctrl control node ids (device)
    const uint32_t *truth;  // n_ctrl x truth_k exact neighbour ids (device)
    uint32_t n_ctrl, truth_k;
    uint32_t *hits;         // iters x n_ctrl (device)
};

// Implemented in wv_nnd.hip.  Writes the finished K-NN graph into the
// index's graph buffers (allocated through index_alloc_graph).  efanna:
// EFANNA's refine, whose pools come from init_ids rows of init_stride ids
// (or InitRandom's when init_ids is null) and whose first join runs over
// 2S random samples per node; else KGraph's, with init_ids rows of S ids.
int nndescent_build(wv_index *ix, const wv_nnd_params *p, const uint32_t *init_ids,
                    const uint32_t *init_cnt, wv_build_stats *stats, uint32_t *graph_out,
                    float *gdist_out, uint32_t *deg_out, bool efanna = false, uint32_t init_stride = 0,
                    const NndMonitor *monitor = nullptr);

// Implemented in wv_capi.hip.  search_collect: greedy searches of Q
// queries (device rows, dp floats each) over the device graph (graph, deg,
// deg_max) from explicit entries (Q x L), writing each query's C smallest
// distinct evaluated keys (Q x C, KEY_EMPTY-padded) and evaluation counts.
int search_collect(wv_index *ix, const uint32_t *graph, const uint32_t *deg, uint32_t deg_max, const float *d_q,
                   uint32_t Q, const uint32_t *d_entries, uint32_t L, uint32_t C, uint64_t *d_full,
                   uint32_t *d_evals);
// Reallocates the index's graph for width W and returns its device buffers.
int index_alloc_graph_dev(wv_index *ix, uint32_t W, uint32_t **ids, float **dist, uint32_t **deg);
// Replaces the index's graph with host arrays (N x W ids / dists, deg[N]).
int index_upload_graph(wv_index *ix, const uint32_t *ids, const float *dist, const uint32_t *deg, uint32_t W);

// How a refiner turns candidates into lists: NSG adds the node's kNN
// neighbours, occludes with alpha 1 and links unreachable nodes; Vamana
// occludes with its alpha and fills lists up to R.
struct RefineRule {
    bool add_knn, fill, conn;
    float alpha;
};
// Implemented in wv_nsg.hip: the NSG (or one Vamana pass) refine of the
// graph knn, kdeg, kstride on the device; the result replaces the index's
// graph.
int nsg_build(wv_index *ix, const wv_nsg_params *p, wv_nsg_stats *stats, const uint32_t *knn, const uint32_t *kdeg,
              uint32_t kstride, const RefineRule &rule);

// Implemented in wv_nsg.hip: the SSG refine of the graph knn (n_try roots
// written to eps_out); the result replaces the index's graph.
int ssg_build(wv_index *ix, const wv_ssg_params *p, uint32_t *eps_out, wv_nsg_stats *stats, const uint32_t *knn,
              const uint32_t *kdeg, uint32_t kstride);
// Implemented in wv_nsg.hip: DPG's prune of the graph knn to `range` a
// node, then the reverse union; the result replaces the index's graph.
int dpg_build(wv_index *ix, uint32_t range, wv_nsg_stats *stats, const uint32_t *knn, const uint32_t *kdeg,
              uint32_t kstride);

}  // namespace wv
