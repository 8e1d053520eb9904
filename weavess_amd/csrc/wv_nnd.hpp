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

// Implemented in wv_nnd.hip.  Writes the finished K-NN graph into the
// index's graph buffers (allocated through index_alloc_graph).  efanna:
// EFANNA's refine, whose pools come from init_ids rows of init_stride ids
// (or InitRandom's when init_ids is null) and whose first join runs over
// 2S random samples per node; else KGraph's, with init_ids rows of S ids.
int nndescent_build(wv_index *ix, const wv_nnd_params *p, const uint32_t *init_ids,
                    const uint32_t *init_cnt, wv_build_stats *stats, uint32_t *graph_out,
                    float *gdist_out, uint32_t *deg_out, bool efanna = false, uint32_t init_stride = 0);

}  // namespace wv
