// index.h -- shared pipeline state (the hot-path members of the reference's
// Index, include/weavess/index.h:1522-1761, and its NNDescent::Neighbor,
// :66-82), plus the device-side handles the HIP stages use.
#pragma once

#include <atomic>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/weavess_hip.h"
#include "parameters.h"
#include "policy.h"

namespace weavess {

// Raised by every HIP stage when the C ABI reports an error (the reference
// calls exit(-1) in the same places; an exception lets callers decide).
struct HipError : std::runtime_error {
    int code;
    HipError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

class Index {
public:
    struct Neighbor {  // include/weavess/index.h:66-82
        unsigned id;
        float distance;
        bool flag;
        Neighbor() = default;
        Neighbor(unsigned i, float d, bool f) : id(i), distance(d), flag(f) {}
        bool operator<(const Neighbor &o) const {
            return distance < o.distance || (distance == o.distance && id < o.id);
        }
    };
    struct SimpleNeighbor {  // :1533-1543
        unsigned id;
        float distance;
        SimpleNeighbor() = default;
        SimpleNeighbor(unsigned i, float d) : id(i), distance(d) {}
        bool operator<(const SimpleNeighbor &o) const {
            return distance < o.distance || (distance == o.distance && id < o.id);
        }
    };
    typedef std::vector<std::vector<SimpleNeighbor>> FinalGraph;
    typedef std::vector<std::vector<unsigned>> LoadGraph;

    explicit Index(int device = 0);
    ~Index();
    Index(const Index &) = delete;
    Index &operator=(const Index &) = delete;

    // data (owned)
    const float *getBaseData() const { return base_.data(); }
    unsigned getBaseLen() const { return base_len_; }
    unsigned getBaseDim() const { return base_dim_; }
    const float *getQueryData() const { return query_.data(); }
    unsigned getQueryLen() const { return query_len_; }
    unsigned getQueryDim() const { return query_dim_; }
    const unsigned *getGroundData() const { return ground_.data(); }
    unsigned getGroundLen() const { return ground_len_; }
    unsigned getGroundDim() const { return ground_dim_; }
    void setBase(std::vector<float> v, unsigned n, unsigned d);
    void setQuery(std::vector<float> v, unsigned n, unsigned d);
    void setGround(std::vector<unsigned> v, unsigned n, unsigned d);

    Parameters &getParam() { return param_; }
    void setParam(const Parameters &p) { param_ = p; }

    unsigned getInitEdgesNum() const { return init_edges_; }          // S
    void setInitEdgesNum(unsigned s) { init_edges_ = s; }
    unsigned getCandidatesEdgesNum() const { return cand_edges_; }    // L
    void setCandidatesEdgesNum(unsigned l) { cand_edges_ = l; }
    unsigned getResultEdgesNum() const { return result_edges_; }      // K
    void setResultEdgesNum(unsigned k) { result_edges_ = k; }

    FinalGraph &getFinalGraph() { return final_graph_; }
    LoadGraph &getLoadGraph() { return load_graph_; }
    LoadGraph &getExactGraph() { return exact_graph_; }

    // navigating (entry) point, ep_ of include/weavess/index.h; set by
    // ComponentRefineEntryCentroidHIP or read from an NSG/Vamana graph file
    unsigned ep_ = 0;
    bool ep_set_ = false;
    // entry set eps_ of include/weavess/index.h, read from an SSG graph file
    std::vector<unsigned> eps_;
    // Vamana's occlusion factor (include/weavess/index.h:215): 1, then 2
    float alpha = 1;
    // the reference's debug flag (include/weavess/index.h:1734): NN-Descent
    // then prints the graph recall of CONTROL_NUM control nodes every round
    bool debug = false;

    // counters (src/component_search_entry.cpp:25, src/component_route.cpp:24,34)
    void addDistCount(uint64_t n = 1) { dist_count_ += n; }
    void addHopCount(uint64_t n = 1) { hop_count_ += n; }
    uint64_t getDistCount() const { return dist_count_; }
    uint64_t getHopCount() const { return hop_count_; }
    void resetDistCount() { dist_count_ = 0; }
    void resetHopCount() { hop_count_ = 0; }

    // device side
    wv_ctx *ctx();                 // created on first use
    wv_index *deviceIndex();       // base uploaded on first use
    void uploadLoadGraph();        // LoadGraph -> device graph (if changed)
    void markLoadGraphChanged() { load_graph_dirty_ = true; }
    void check(int rc) const;      // throws HipError with wv_last_error

    uint64_t seed() const { return param_.has("seed") ? param_.get<uint64_t>("seed") : 0; }

private:
    int device_;
    std::vector<float> base_, query_;
    std::vector<unsigned> ground_;
    unsigned base_len_ = 0, base_dim_ = 0, query_len_ = 0, query_dim_ = 0;
    unsigned ground_len_ = 0, ground_dim_ = 0;
    Parameters param_;
    unsigned init_edges_ = 0, cand_edges_ = 0, result_edges_ = 0;
    FinalGraph final_graph_;
    LoadGraph load_graph_, exact_graph_;
    std::atomic<uint64_t> dist_count_{0}, hop_count_{0};
    wv_ctx *ctx_ = nullptr;
    wv_index *dix_ = nullptr;
    bool load_graph_dirty_ = true;
};

}  // namespace weavess
