// policy.h -- stage selectors of the pipeline (include/weavess/policy.h:9-33
// of the reference, hot-path values only, plus the HIP stages).
#pragma once

namespace weavess {

enum TYPE {
    INDEX_KGRAPH,
    INDEX_EXACT_KNNG,

    INIT_RANDOM,            // ComponentInitRandom (host, reference semantics)
    INIT_RANDOM_HIP,        // random S-NN initialisation on the GPU

    REFINE_NN_DESCENT,      // accepted: runs the HIP NN-Descent (the only refine here)
    REFINE_NN_DESCENT_HIP,  // NN-Descent on the GPU, FP32 join
    REFINE_NN_DESCENT_HIP_BF16,
    REFINE_NN_DESCENT_HIP_FP8,

    SEARCH_ENTRY_RAND,      // L random entries per query (GenRandom)

    L_SEARCH_ASCEND, L_SEARCH_SET_RECALL, L_SEARCH_ASSIGN,

    ROUTER_GREEDY,          // accepted: runs the HIP greedy router
    ROUTER_GREEDY_HIP,
};

}  // namespace weavess
