// component.h -- the pipeline's plug-in stages.
//
// Stage bases keep the reference's virtual hooks and signatures
// (include/weavess/component.h:12-37, 361-366, 752-757, 827-832).  Unlike
// the reference, ~Component does not delete the shared Index (component.h:16
// there deletes it, which is why its builder leaks every component); the
// Index is owned by IndexBuilder alone and device resources by the Index.
#pragma once

#include <vector>

#include "index.h"

namespace weavess {

class Component {
public:
    explicit Component(Index *index) : index(index) {}
    virtual ~Component() = default;

protected:
    Index *index = nullptr;
};

class ComponentLoad : public Component {
public:
    explicit ComponentLoad(Index *index) : Component(index) {}
    // src/component_load.cpp:67-104
    virtual void LoadInner(const char *data_file, const char *query_file, const char *ground_file,
                           Parameters &parameters);
};

class ComponentInit : public Component {
public:
    explicit ComponentInit(Index *index) : Component(index) {}
    virtual void InitInner() = 0;
};

// Random S-NN initialisation on the GPU (src/component_init.cpp:11-45):
// FinalGraph[i] = S random distinct ids != i, ascending (dist, id).
// InitRandom (src/component_init.cpp:11-68) on the GPU: "S" random
// neighbours a node (GenRandom, self skipped), sorted, into FinalGraph.
// With rand_init, ComponentInitRand (:979-1020): the count is "L".
class ComponentInitRandomHIP : public ComponentInit {
public:
    explicit ComponentInitRandomHIP(Index *index, bool rand_init = false)
        : ComponentInit(index), rand_init_(rand_init) {}
    void InitInner() override;

private:
    bool rand_init_;
};

class ComponentRefine : public Component {
public:
    explicit ComponentRefine(Index *index) : Component(index) {}
    virtual void RefineInner() = 0;
};

// NN-Descent + PruneNaive on the GPU (src/component_refine.cpp:12-239,
// src/component_prune.cpp:8-18).  Reads "L","K","R","ITER" like
// ComponentRefineNNDescent::SetConfigs (:79-85), S from the init stage and
// the initial lists from FinalGraph; writes FinalGraph (<= K per node).
// efanna: ComponentRefineEFANNA (:959-1183) instead, which reads S from the
// parameters (:1028-1034), takes whole FinalGraph rows as initial pools (a
// row longer than L is cut to its L nearest) and joins 2S random samples
// per node in its first round.
class ComponentRefineNNDescentHIP : public ComponentRefine {
public:
    ComponentRefineNNDescentHIP(Index *index, unsigned dtype = WV_DTYPE_F32, bool efanna = false)
        : ComponentRefine(index), dtype_(dtype), efanna_(efanna) {}
    void RefineInner() override;
    const wv_build_stats &stats() const { return stats_; }

private:
    unsigned dtype_;
    bool efanna_;
    wv_build_stats stats_{};
};

// ComponentRefineNSG (src/component_refine.cpp:393-436) on the GPU: reads
// "L_refine", "R_refine", "C_refine"; finds ep_ with
// ComponentRefineEntryCentroidHIP unless it is set; refines FinalGraph (the
// kNN graph) into the NSG graph, connectivity links included.
class ComponentRefineNSGHIP : public ComponentRefine {
public:
    explicit ComponentRefineNSGHIP(Index *index) : ComponentRefine(index) {}
    void RefineInner() override;
    const wv_nsg_stats &stats() const { return stats_; }

private:
    wv_nsg_stats stats_{};
};

// ComponentRefineVAMANA (src/component_refine.cpp:794-958) on the GPU, one
// pass: reads "L_refine", "R_refine" (and "C_refine" when set, else 1024
// candidates); ep_ from ComponentRefineEntryCentroidHIP unless set; the
// Index's alpha (1 on the first pass; the pass sets it to 2, :871).
class ComponentRefineVamanaHIP : public ComponentRefine {
public:
    explicit ComponentRefineVamanaHIP(Index *index) : ComponentRefine(index) {}
    void RefineInner() override;
    const wv_nsg_stats &stats() const { return stats_; }

private:
    wv_nsg_stats stats_{};
};

// ComponentRefineDPG (src/component_refine.cpp:728-790) on the GPU: keeps
// "K" / 2 a node (L_dpg, :763) by occlusion counts, then mirrors every edge
// (ComponentConnReverse).
class ComponentRefineDPGHIP : public ComponentRefine {
public:
    explicit ComponentRefineDPGHIP(Index *index) : ComponentRefine(index) {}
    void RefineInner() override;

private:
    wv_nsg_stats stats_{};
};

// ComponentRefineSSG (src/component_refine.cpp:550-727) on the GPU: reads
// "L_refine", "R_refine", "A", "n_try" (SetConfigs, :584-591); writes the
// SSG graph to FinalGraph and its connectivity roots to eps_.
class ComponentRefineSSGHIP : public ComponentRefine {
public:
    explicit ComponentRefineSSGHIP(Index *index) : ComponentRefine(index) {}
    void RefineInner() override;

private:
    wv_nsg_stats stats_{};
};

class ComponentSearchEntry : public Component {
public:
    explicit ComponentSearchEntry(Index *index) : Component(index) {}
    virtual void SearchEntryInner(unsigned query, std::vector<Index::Neighbor> &pool) = 0;
};

// SearchEntryRand (src/component_search_entry.cpp:9-30): pool sized L+1,
// the first L slots hold GenRandom ids (same seeded stream the device uses),
// flagged.  Distances are left 0 here: the GPU router computes them.
class ComponentSearchEntryRand : public ComponentSearchEntry {
public:
    explicit ComponentSearchEntryRand(Index *index) : ComponentSearchEntry(index) {}
    void SearchEntryInner(unsigned query, std::vector<Index::Neighbor> &pool) override;
};

// ComponentRefineEntryCentroid (src/component_entry.cpp:9-28): ep_ = the
// node nearest to the mean of the base.  The reference finds it with an
// approximate greedy search from a random node; here it is the exact FP64
// nearest neighbour of the centroid, found on the GPU (the limit of that
// search).  The centroid is summed in double and rounded to float.
class ComponentRefineEntryCentroidHIP : public Component {
public:
    explicit ComponentRefineEntryCentroidHIP(Index *index) : Component(index) {}
    void EntryInner();
};

// SearchEntryCentroid (src/component_search_entry.cpp:33-67): the first
// min(L, degree) neighbours of ep_ in the load graph, then distinct random
// ids until L.  The reference draws the fill with rand(); here it comes from
// the query's counter stream (draw t mod N, t = 0, 1, ..., skipping ids
// already taken).
class ComponentSearchEntryCentroid : public ComponentSearchEntry {
public:
    explicit ComponentSearchEntryCentroid(Index *index) : ComponentSearchEntry(index) {}
    void SearchEntryInner(unsigned query, std::vector<Index::Neighbor> &pool) override;
};

// SearchEntrySubCentroid (src/component_search_entry.cpp:70-99): the entry
// set eps_ first, then random ids until L.  The reference overwrites the
// head of a GenRandom set with eps_, so an id can appear twice, and it
// reads each base row one float late (its `x++`); here the ids are
// distinct (eps_ deduplicated, the fill drawn like the centroid entry's
// from its own stream) and distances are plain.  eps_ longer than L is
// rejected, as the reference's assert does.
class ComponentSearchEntrySubCentroid : public ComponentSearchEntry {
public:
    explicit ComponentSearchEntrySubCentroid(Index *index) : ComponentSearchEntry(index) {}
    void SearchEntryInner(unsigned query, std::vector<Index::Neighbor> &pool) override;
};

class ComponentSearchRoute : public Component {
public:
    explicit ComponentSearchRoute(Index *index) : Component(index) {}
    virtual void RouteInner(unsigned query, std::vector<Index::Neighbor> &pool,
                            std::vector<unsigned> &res) = 0;
    // Whole query set at once; the default loops over RouteInner.
    virtual void RouteBatch(unsigned nq, const std::vector<std::vector<Index::Neighbor>> *pools,
                            std::vector<std::vector<unsigned>> &res);
};

// Greedy best-first search on the GPU (src/component_route.cpp:9-53).
// RouteInner takes its entries from `pool` and writes the final sorted pool
// back into it; RouteBatch searches every query in one launch, with the
// entries of `pools` when given, else with entries generated on the device
// from the same seeded stream as ComponentSearchEntryRand.
class ComponentSearchRouteGreedyHIP : public ComponentSearchRoute {
public:
    explicit ComponentSearchRouteGreedyHIP(Index *index) : ComponentSearchRoute(index) {}
    void RouteInner(unsigned query, std::vector<Index::Neighbor> &pool,
                    std::vector<unsigned> &res) override;
    void RouteBatch(unsigned nq, const std::vector<std::vector<Index::Neighbor>> *pools,
                    std::vector<std::vector<unsigned>> &res) override;
    double last_kernel_ms() const { return kernel_ms_; }

private:
    double kernel_ms_ = 0.0;
};

// GenRandom (include/weavess/util.h:13-27) on the counter-based stream of
// (seed, tag, idx) that the device kernels use.
void GenRandomStream(uint64_t seed, uint64_t tag, uint64_t idx, unsigned *addr, unsigned size,
                     unsigned N);

}  // namespace weavess
