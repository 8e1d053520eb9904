// builder.h -- chained pipeline driver (include/weavess/builder.h:11-59).
#pragma once

#include <chrono>

#include "component.h"

namespace weavess {

class IndexBuilder {
public:
    // num_threads is kept for interface parity (the reference sets OpenMP
    // threads, builder.h:13-16); the stages here run on `device`.
    explicit IndexBuilder(unsigned num_threads = 1, int device = 0);
    virtual ~IndexBuilder();

    IndexBuilder *load(const char *data_file, const char *query_file, const char *ground_file,
                       Parameters &parameters);
    IndexBuilder *init(TYPE type, bool debug = false);
    IndexBuilder *refine(TYPE type, bool debug = false);
    IndexBuilder *search(TYPE entry_type, TYPE route_type, TYPE L_type);
    IndexBuilder *save_graph(TYPE type, const char *graph_file);
    IndexBuilder *load_graph(TYPE type, const char *graph_file);
    IndexBuilder *print_index_info(TYPE type);

    double graph_quality(TYPE type);
    std::chrono::duration<double> GetBuildTime() { return e_ - s_; }
    void peak_memory_footprint();

    Index *index() { return index_; }
    // last search: recall@K_search, QPS, L used
    double last_recall() const { return last_recall_; }
    double last_qps() const { return last_qps_; }
    unsigned last_L() const { return last_L_; }
    const std::vector<std::vector<unsigned>> &last_results() const { return res_; }

private:
    double run_queries(ComponentSearchEntry *a, ComponentSearchRoute *b, bool batch, double &secs);
    double accuracy(unsigned K);

    Index *index_;
    std::chrono::high_resolution_clock::time_point s_, e_;
    std::vector<std::vector<unsigned>> res_;
    double last_recall_ = 0.0, last_qps_ = 0.0;
    unsigned last_L_ = 0;
};

}  // namespace weavess
