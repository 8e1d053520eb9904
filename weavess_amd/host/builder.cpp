// builder.cpp -- IndexBuilder: stage dispatch, search harness, graph I/O
// and metrics (src/builder.cpp:22-487, 714-778, 1068-1076, 1301-1308).
#include "weavess/builder.h"

#include <algorithm>
#include <fstream>
#include <iostream>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>

namespace weavess {

IndexBuilder::IndexBuilder(unsigned /*num_threads*/, int device) : index_(new Index(device)) {}

IndexBuilder::~IndexBuilder() { delete index_; }

IndexBuilder *IndexBuilder::load(const char *data_file, const char *query_file, const char *ground_file,
                                 Parameters &parameters) {
    ComponentLoad(index_).LoadInner(data_file, query_file, ground_file, parameters);
    std::cout << "base data len : " << index_->getBaseLen() << "\nbase data dim : " << index_->getBaseDim()
              << "\nquery data len : " << index_->getQueryLen() << "\nground truth data len : "
              << index_->getGroundLen() << "\nground truth data dim : " << index_->getGroundDim() << std::endl;
    return this;
}

IndexBuilder *IndexBuilder::init(TYPE type, bool) {
    s_ = std::chrono::high_resolution_clock::now();
    std::unique_ptr<ComponentInit> a;
    if (type == INIT_RANDOM || type == INIT_RANDOM_HIP) {
        std::cout << "__INIT : RANDOM (HIP)__" << std::endl;
        a.reset(new ComponentInitRandomHIP(index_));
    } else if (type == INIT_RAND) {
        std::cout << "__INIT : RAND (HIP)__" << std::endl;  // src/builder.cpp:51-53
        a.reset(new ComponentInitRandomHIP(index_, true));
    } else {
        throw std::invalid_argument("__INIT : WRONG TYPE__");  // src/builder.cpp:91-93
    }
    a->InitInner();
    e_ = std::chrono::high_resolution_clock::now();
    return this;
}

IndexBuilder *IndexBuilder::refine(TYPE type, bool debug) {
    std::unique_ptr<ComponentRefine> a;
    index_->debug = debug;
    if (type == REFINE_NN_DESCENT || type == REFINE_NN_DESCENT_HIP) {
        std::cout << "__REFINE : KGRAPH (HIP)__" << std::endl;
        a.reset(new ComponentRefineNNDescentHIP(index_, WV_DTYPE_F32));
    } else if (type == REFINE_NN_DESCENT_HIP_BF16) {
        a.reset(new ComponentRefineNNDescentHIP(index_, WV_DTYPE_BF16));
    } else if (type == REFINE_NN_DESCENT_HIP_FP8) {
        a.reset(new ComponentRefineNNDescentHIP(index_, WV_DTYPE_FP8));
    } else if (type == REFINE_SSG) {
        std::cout << "__REFINE : SSG (HIP)__" << std::endl;  // src/builder.cpp:128-130
        a.reset(new ComponentRefineSSGHIP(index_));
    } else if (type == REFINE_DPG) {
        std::cout << "__REFINE : DPG (HIP)__" << std::endl;  // src/builder.cpp:131-133
        a.reset(new ComponentRefineDPGHIP(index_));
    } else if (type == REFINE_VAMANA) {
        std::cout << "__REFINE : VAMANA (HIP)__" << std::endl;  // src/builder.cpp:134-136
        a.reset(new ComponentRefineVamanaHIP(index_));
    } else if (type == REFINE_NSG) {
        std::cout << "__REFINE : NSG (HIP)__" << std::endl;  // src/builder.cpp:125-127
        a.reset(new ComponentRefineNSGHIP(index_));
    } else if (type == REFINE_EFANNA || type == REFINE_EFANNA_HIP_BF16 || type == REFINE_EFANNA_HIP_FP8) {
        std::cout << "__REFINE : EFANNA (HIP)__" << std::endl;  // src/builder.cpp:137-139
        a.reset(new ComponentRefineNNDescentHIP(
            index_, type == REFINE_EFANNA ? WV_DTYPE_F32 : type == REFINE_EFANNA_HIP_BF16 ? WV_DTYPE_BF16 : WV_DTYPE_FP8,
            true));
    } else {
        throw std::invalid_argument("__REFINE : WRONG TYPE__");
    }
    a->RefineInner();
    e_ = std::chrono::high_resolution_clock::now();
    return this;
}

double IndexBuilder::accuracy(unsigned K) {
    // src/builder.cpp:465-480
    const unsigned gl = index_->getGroundLen(), gd = index_->getGroundDim();
    if (gl == 0) return 0.0;
    if (gd < K) throw std::invalid_argument("ground truth narrower than K_search");
    uint64_t miss = 0;
    for (unsigned i = 0; i < gl && i < res_.size(); ++i) {
        if (res_[i].empty()) continue;
        for (unsigned j = 0; j < K; ++j) {
            unsigned k = 0;
            for (; k < K; ++k)
                if (res_[i][j] == index_->getGroundData()[(size_t)i * gd + k]) break;
            if (k == K) ++miss;
        }
    }
    return 1.0 - (double)miss / ((double)gl * K);
}

double IndexBuilder::run_queries(ComponentSearchEntry *a, ComponentSearchRoute *b, bool batch, double &secs) {
    const unsigned nq = index_->getQueryLen();
    auto t0 = std::chrono::high_resolution_clock::now();
    if (batch && dynamic_cast<ComponentSearchEntryRand *>(a)) {
        b->RouteBatch(nq, nullptr, res_);  // entries generated on the device
    } else if (batch) {
        std::vector<std::vector<Index::Neighbor>> pools(nq);
        for (unsigned i = 0; i < nq; ++i) a->SearchEntryInner(i, pools[i]);
        b->RouteBatch(nq, &pools, res_);
    } else {
        res_.assign(nq, {});
        for (unsigned i = 0; i < nq; ++i) {
            std::vector<Index::Neighbor> pool;
            a->SearchEntryInner(i, pool);
            b->RouteInner(i, pool, res_[i]);
        }
    }
    auto t1 = std::chrono::high_resolution_clock::now();
    secs = std::chrono::duration<double>(t1 - t0).count();
    return secs;
}

IndexBuilder *IndexBuilder::search(TYPE entry_type, TYPE route_type, TYPE L_type) {
    std::cout << "__SEARCH__" << std::endl;
    const unsigned K = 10;  // src/builder.cpp:190-192
    index_->getParam().set<unsigned>("K_search", K);
    if (entry_type != SEARCH_ENTRY_RAND && entry_type != SEARCH_ENTRY_CENTROID &&
        entry_type != SEARCH_ENTRY_SUB_CENTROID)
        throw std::invalid_argument("__SEARCH ENTRY : WRONG TYPE__");
    if (route_type != ROUTER_GREEDY && route_type != ROUTER_GREEDY_HIP)
        throw std::invalid_argument("__ROUTER : WRONG TYPE__");
    if (index_->getLoadGraph().empty() && !index_->getFinalGraph().empty()) {
        // graph built in this process: search it without a file round trip
        auto &lg = index_->getLoadGraph();
        lg.clear();
        for (auto &row : index_->getFinalGraph()) {
            lg.emplace_back();
            for (auto &nb : row) lg.back().push_back(nb.id);
        }
        index_->markLoadGraphChanged();
    }
    ComponentSearchEntryRand a_rand(index_);
    ComponentSearchEntryCentroid a_cent(index_);
    ComponentSearchEntrySubCentroid a_sub(index_);
    ComponentSearchEntry &a = entry_type == SEARCH_ENTRY_CENTROID       ? static_cast<ComponentSearchEntry &>(a_cent)
                              : entry_type == SEARCH_ENTRY_SUB_CENTROID ? static_cast<ComponentSearchEntry &>(a_sub)
                                                                        : static_cast<ComponentSearchEntry &>(a_rand);
    if (entry_type == SEARCH_ENTRY_CENTROID && !index_->ep_set_) {
        std::cout << "__ENTRY : CENTROID__" << std::endl;
        ComponentRefineEntryCentroidHIP(index_).EntryInner();
    }
    ComponentSearchRouteGreedyHIP b(index_);
    const bool batch = !index_->getParam().has("per_query") || index_->getParam().get<unsigned>("per_query") == 0;
    const unsigned nq = index_->getQueryLen();

    auto one = [&](unsigned L) {
        if (L < K) throw std::invalid_argument("search_L cannot be smaller than search_K!");
        index_->getParam().set<unsigned>("L_search", L);
        std::cout << "SEARCH_L : " << L << std::endl;
        double secs = 0;
        run_queries(&a, &b, batch, secs);
        std::cout << "search time: " << secs << "\nDistCount: " << index_->getDistCount()
                  << "\nHopCount: " << index_->getHopCount() << std::endl;
        index_->resetDistCount();
        index_->resetHopCount();
        const double acc = accuracy(K);
        std::cout << K << " NN accuracy: " << acc << std::endl;
        last_recall_ = acc;
        last_qps_ = secs > 0 ? nq / secs : 0.0;
        last_L_ = L;
        return acc;
    };

    if (L_type == L_SEARCH_ASSIGN) {
        one(index_->getParam().get<unsigned>("L_search"));
    } else if (L_type == L_SEARCH_ASCEND) {  // src/builder.cpp:354-430
        unsigned a1 = 5, a2 = 8;
        for (unsigned i = 0; i < 10; ++i) {
            const unsigned L = a1 + a2;
            a1 = a2;
            a2 = L;
            if (L >= index_->getBaseLen()) break;
            one(L);
        }
    } else if (L_type == L_SEARCH_SET_RECALL) {  // src/builder.cpp:262-353
        const float acc_set = index_->getParam().has("acc_set") ? index_->getParam().get<float>("acc_set") : 0.99f;
        const unsigned sg = 1000;
        std::set<unsigned> visited;
        bool flag = false;
        int L_sl = 1;
        unsigned L = K, L_min = 0x7fffffff;
        visited.insert(L);
        while (true) {
            const double acc = one(L);
            if (acc_set - acc <= 0) {
                if (L_min > L) L_min = L;
                if (L == K || L_sl == 1) break;
                if (!flag) {
                    L_sl < 0 ? L_sl-- : L_sl++;
                    flag = true;
                }
                L_sl /= 2;
                if (L_sl == 0) break;
                if (L_sl > 0) L_sl = -L_sl;
            } else {
                if (L_min < L) break;
                L_sl = (int)(sg * (acc_set - acc));
                if (L_sl == 0) L_sl++;
                flag = false;
            }
            const long nl = (long)L + L_sl;
            if (nl < (long)K || nl >= (long)index_->getBaseLen()) break;
            L = (unsigned)nl;
            if (visited.count(L)) break;
            visited.insert(L);
        }
        std::cout << "L_min: " << L_min << std::endl;
        if (L_min != 0x7fffffffu) one(L_min);  // leave last_* at the minimal L
    } else {
        throw std::invalid_argument("__L TYPE : WRONG TYPE__");
    }
    return this;
}

IndexBuilder *IndexBuilder::save_graph(TYPE type, const char *graph_file) {
    if (type != INDEX_KGRAPH && type != INDEX_NSG && type != INDEX_VAMANA && type != INDEX_SSG)
        throw std::invalid_argument("graph type not written here");
    std::ofstream out(graph_file, std::ios::binary);
    if (!out.is_open()) throw std::runtime_error(std::string("cannot write ") + graph_file);
    if (type == INDEX_NSG || type == INDEX_VAMANA) {  // src/builder.cpp:920-921
        if (!index_->ep_set_) throw std::invalid_argument("no entry point to write");
        out.write(reinterpret_cast<const char *>(&index_->ep_), sizeof(unsigned));
    } else if (type == INDEX_SSG) {  // src/builder.cpp:922-925
        const unsigned n_ep = (unsigned)index_->eps_.size();
        out.write(reinterpret_cast<const char *>(&n_ep), sizeof n_ep);
        out.write(reinterpret_cast<const char *>(index_->eps_.data()), (std::streamsize)n_ep * sizeof(unsigned));
    }
    for (const auto &row : index_->getFinalGraph()) {  // src/builder.cpp:1068-1076
        const unsigned GK = (unsigned)row.size();
        out.write(reinterpret_cast<const char *>(&GK), sizeof GK);
        for (const auto &nb : row) out.write(reinterpret_cast<const char *>(&nb.id), sizeof nb.id);
    }
    if (!out) throw std::runtime_error("graph write failed");
    return this;
}

IndexBuilder *IndexBuilder::load_graph(TYPE type, const char *graph_file) {
    if (type != INDEX_KGRAPH && type != INDEX_EXACT_KNNG && type != INDEX_NSG && type != INDEX_VAMANA &&
        type != INDEX_SSG)
        throw std::invalid_argument("graph type not read here");
    std::ifstream in(graph_file, std::ios::binary);
    if (!in.is_open()) throw std::runtime_error(std::string("load graph error: ") + graph_file);
    if (type == INDEX_NSG || type == INDEX_VAMANA) {  // src/builder.cpp:1089-1090
        unsigned ep = 0;
        in.read(reinterpret_cast<char *>(&ep), sizeof ep);
        if (!in) throw std::runtime_error("truncated graph header");
        index_->ep_ = ep;
        index_->ep_set_ = true;
    } else if (type == INDEX_SSG) {  // src/builder.cpp:1091-1095
        unsigned n_ep = 0;
        in.read(reinterpret_cast<char *>(&n_ep), sizeof n_ep);
        if (!in || n_ep > index_->getBaseLen()) throw std::runtime_error("bad graph header");
        index_->eps_.resize(n_ep);
        in.read(reinterpret_cast<char *>(index_->eps_.data()), (std::streamsize)n_ep * sizeof(unsigned));
        if (!in) throw std::runtime_error("truncated graph header");
    }
    auto &dst = type == INDEX_EXACT_KNNG ? index_->getExactGraph() : index_->getLoadGraph();
    dst.clear();
    while (true) {  // src/builder.cpp:1301-1308
        unsigned GK;
        in.read(reinterpret_cast<char *>(&GK), sizeof GK);
        if (in.eof()) break;
        std::vector<unsigned> tmp(GK);
        in.read(reinterpret_cast<char *>(tmp.data()), (std::streamsize)GK * sizeof(unsigned));
        if (!in) throw std::runtime_error("truncated graph file");
        dst.push_back(std::move(tmp));
    }
    if (type != INDEX_EXACT_KNNG) index_->markLoadGraphChanged();
    return this;
}

double IndexBuilder::graph_quality(TYPE) {
    // src/builder.cpp:714-778 with the self entry skipped wherever it is
    if (index_->getExactGraph().empty()) {
        const std::string p = index_->getParam().get<std::string>("exact_knng_path");
        load_graph(INDEX_EXACT_KNNG, p.c_str());
    }
    auto &G = index_->getLoadGraph();
    auto &E = index_->getExactGraph();
    const unsigned control = 10;
    double mean = 0.0;
    const size_t n = std::min(G.size(), E.size());
    for (size_t i = 0; i < n; ++i) {
        std::vector<unsigned> v;
        for (unsigned x : E[i])
            if (x != i) v.push_back(x);
        unsigned q = 0;
        for (unsigned j = 0; j < control && j < v.size(); ++j)
            if (std::find(G[i].begin(), G[i].end(), v[j]) != G[i].end()) ++q;
        const unsigned den = std::min<unsigned>(std::min<unsigned>(control, (unsigned)G[i].size()), (unsigned)v.size());
        if (den) mean += (double)q / den;
    }
    const double gq = n ? mean / n : 0.0;
    std::cout << "Graph Quality : " << gq << std::endl;
    return gq;
}

IndexBuilder *IndexBuilder::print_index_info(TYPE type) {
    auto &G = index_->getLoadGraph();
    size_t edges = 0, mx = 0, mn = G.empty() ? 0 : SIZE_MAX;
    for (auto &r : G) {
        edges += r.size();
        mx = std::max(mx, r.size());
        mn = std::min(mn, r.size());
    }
    std::cout << "nodes: " << G.size() << " edges: " << edges << " max degree: " << mx << " min degree: " << mn
              << std::endl;
    if ((type == INDEX_NSG || type == INDEX_VAMANA) && index_->ep_set_) std::cout << "ep: " << index_->ep_ << std::endl;
    if (type == INDEX_SSG) std::cout << "eps: " << index_->eps_.size() << std::endl;
    if (index_->getParam().has("exact_knng_path") || !index_->getExactGraph().empty()) graph_quality(type);
    return this;
}

void IndexBuilder::peak_memory_footprint() {
    std::ifstream f("/proc/self/status");
    std::string line;
    while (std::getline(f, line))
        if (line.rfind("VmPeak", 0) == 0 || line.rfind("VmHWM", 0) == 0) std::cout << line << std::endl;
}

}  // namespace weavess
