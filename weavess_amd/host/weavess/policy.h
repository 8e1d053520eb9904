// policy.h -- stage selectors of the pipeline (include/weavess/policy.h:9-33
// of the reference, hot-path values only, plus the HIP stages).
#pragma once

namespace weavess {

enum TYPE {
    INDEX_KGRAPH,
    INDEX_EXACT_KNNG,
    INDEX_NSG,              // graph file headed by ep_ (src/builder.cpp:920-921)
    INDEX_VAMANA,           // same layout as INDEX_NSG
    INDEX_SSG,              // graph file headed by n_ep and eps_ (src/builder.cpp:922-925)

    INIT_RANDOM,            // ComponentInitRandom (host, reference semantics)
    INIT_RANDOM_HIP,        // random S-NN initialisation on the GPU
    INIT_RAND,              // ComponentInitRand: "L" random neighbours a node, on the GPU

    REFINE_NN_DESCENT,      // accepted: runs the HIP NN-Descent (the only refine here)
    REFINE_NN_DESCENT_HIP,  // NN-Descent on the GPU, FP32 join
    REFINE_NN_DESCENT_HIP_BF16,
    REFINE_NN_DESCENT_HIP_FP8,
    REFINE_EFANNA,          // EFANNA's NN-Descent on the GPU, FP32 join
    REFINE_EFANNA_HIP_BF16,
    REFINE_EFANNA_HIP_FP8,
    REFINE_NSG,             // NSG refine of the kNN graph on the GPU
    REFINE_VAMANA,          // one Vamana pass on the GPU (alpha 1, then 2 on the next call)
    REFINE_DPG,             // DPG refine of the kNN graph on the GPU
    REFINE_SSG,             // SSG refine of the kNN graph on the GPU (sets eps_)

    SEARCH_ENTRY_RAND,      // L random entries per query (GenRandom)
    SEARCH_ENTRY_CENTROID,  // the entry point's neighbours, then distinct random ids
    SEARCH_ENTRY_SUB_CENTROID,  // the entry set eps_, then distinct random ids

    L_SEARCH_ASCEND, L_SEARCH_SET_RECALL, L_SEARCH_ASSIGN,

    ROUTER_GREEDY,          // accepted: runs the HIP greedy router
    ROUTER_GREEDY_HIP,
};

}  // namespace weavess
