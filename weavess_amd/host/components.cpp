// components.cpp -- the HIP stages behind the reference's plug-in hooks.
#include <algorithm>
#include <fstream>
#include <iostream>
#include <stdexcept>

#include "weavess/component.h"

namespace weavess {

namespace {
inline uint64_t fmix64(uint64_t z) {  // splitmix64 finaliser
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    return z;
}
constexpr uint64_t GOLDEN = 0x9E3779B97F4A7C15ull;
constexpr uint64_t TAG_ENTRY = 0x454E5452ull;
constexpr uint64_t TAG_CONTROL = 0x4354524Cull;     // "CTRL": the debug monitor's control nodes
constexpr unsigned CONTROL_NUM = 100;               // include/weavess/index.h:13
constexpr uint64_t TAG_CENTROID = 0x43454E54ull;     // "CENT": centroid-entry fill
constexpr uint64_t TAG_SUBCENTROID = 0x53554243ull;  // "SUBC": sub-centroid-entry fill
uint64_t stream_of(uint64_t seed, uint64_t tag, uint64_t idx) {
    uint64_t s = fmix64(seed + GOLDEN);
    s = fmix64(s ^ (tag * GOLDEN));
    return fmix64(s ^ (idx + 0x632BE59BD9B4E019ull));
}

// L distinct entry ids: the ids of `head` (below N, first occurrence, at
// most L of them), then draws t mod N, t = 0, 1, ..., of the query's
// counter stream, skipping ids already taken.
void head_then_fill(const std::vector<unsigned> &head, uint64_t stream, unsigned L, unsigned N,
                    std::vector<Index::Neighbor> &pool) {
    if (L > N) throw std::invalid_argument("L_search above the base size");
    std::vector<unsigned> ids;
    ids.reserve(L);
    std::vector<bool> taken(N, false);
    for (unsigned nb : head) {
        if (ids.size() == L) break;
        if (nb < N && !taken[nb]) {
            taken[nb] = true;
            ids.push_back(nb);
        }
    }
    for (uint64_t t = 0; ids.size() < L; ++t) {
        const unsigned id = (unsigned)((uint32_t)(fmix64(stream + (t + 1) * GOLDEN) >> 32) % N);
        if (taken[id]) continue;
        taken[id] = true;
        ids.push_back(id);
    }
    pool.resize(L + 1);
    for (unsigned i = 0; i < L; ++i) pool[i] = Index::Neighbor(ids[i], 0.0f, true);
}
}  // namespace

void GenRandomStream(uint64_t seed, uint64_t tag, uint64_t idx, unsigned *addr, unsigned size, unsigned N) {
    if (size >= N) throw std::invalid_argument("GenRandom needs size < N");
    const uint64_t s = stream_of(seed, tag, idx);
    auto draw = [s](uint64_t k) { return (uint32_t)(fmix64(s + (k + 1) * GOLDEN) >> 32); };
    for (unsigned i = 0; i < size; ++i) addr[i] = draw(i) % (N - size);
    std::sort(addr, addr + size);
    for (unsigned i = 1; i < size; ++i)
        if (addr[i] <= addr[i - 1]) addr[i] = addr[i - 1] + 1;
    const uint64_t off = draw(size) % N;
    for (unsigned i = 0; i < size; ++i) addr[i] = (unsigned)(((uint64_t)addr[i] + off) % N);
}

// ------------------------------------------------------------------ load
namespace {
// fvecs/ivecs: per row an int32 dimension then dim 4-byte values
// (src/component_load.cpp:9-28).
template <class T>
void load_vecs(const char *path, std::vector<T> &data, unsigned &num, unsigned &dim) {
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) throw std::runtime_error(std::string("open file error: ") + path);
    int32_t d = 0;
    in.read(reinterpret_cast<char *>(&d), 4);
    if (!in || d <= 0) throw std::runtime_error(std::string("bad vecs header: ") + path);
    in.seekg(0, std::ios::end);
    const size_t bytes = (size_t)in.tellg();
    dim = (unsigned)d;
    num = (unsigned)(bytes / ((size_t)dim + 1) / 4);
    data.resize((size_t)num * dim);
    in.seekg(0, std::ios::beg);
    for (size_t i = 0; i < num; ++i) {
        in.seekg(4, std::ios::cur);
        in.read(reinterpret_cast<char *>(data.data() + i * dim), (std::streamsize)dim * sizeof(T));
    }
    if (!in) throw std::runtime_error(std::string("short vecs file: ") + path);
}
}  // namespace

void ComponentLoad::LoadInner(const char *data_file, const char *query_file, const char *ground_file,
                              Parameters &parameters) {
    std::vector<float> b, q;
    std::vector<unsigned> g;
    unsigned n, d, qn, qd, gn = 0, gd = 0;
    load_vecs(data_file, b, n, d);
    load_vecs(query_file, q, qn, qd);
    if (qd != d) throw std::runtime_error("query dimension differs from base dimension");
    if (ground_file && *ground_file) load_vecs(ground_file, g, gn, gd);
    index->setBase(std::move(b), n, d);
    index->setQuery(std::move(q), qn, qd);
    index->setGround(std::move(g), gn, gd);
    index->setParam(parameters);
}

// ------------------------------------------------------------------ init / refine
void ComponentInitRandomHIP::InitInner() {
    // component_init.cpp:47-49 ("S"), or ComponentInitRand's "L" (:1027-1029)
    const unsigned S = index->getParam().get<unsigned>(rand_init_ ? "L" : "S");
    if (!rand_init_) index->setInitEdgesNum(S);
    wv_nnd_params p{};
    p.K = S;
    p.L = S;
    p.S = S;
    p.R = 0;
    p.iters = 0;  // InitRandom only
    p.dtype = WV_DTYPE_F32;
    p.seed = index->seed();
    wv_index *ix = index->deviceIndex();
    index->check(wv_index_build_nndescent(ix, &p, nullptr, nullptr, nullptr));
    const unsigned N = index->getBaseLen();
    std::vector<uint32_t> ids((size_t)N * S), deg(N);
    std::vector<float> dist((size_t)N * S);
    index->check(wv_index_get_graph(ix, ids.data(), dist.data(), deg.data(), nullptr));
    auto &fg = index->getFinalGraph();
    fg.assign(N, {});
    for (unsigned i = 0; i < N; ++i) {
        fg[i].reserve(S);
        for (unsigned j = 0; j < deg[i]; ++j)
            fg[i].emplace_back(ids[(size_t)i * S + j], dist[(size_t)i * S + j]);
    }
}

void ComponentRefineNNDescentHIP::RefineInner() {
    auto &P = index->getParam();
    // SetConfigs, component_refine.cpp:79-85
    index->setCandidatesEdgesNum(P.get<unsigned>("L"));
    index->setResultEdgesNum(P.get<unsigned>("K"));
    wv_nnd_params p{};
    p.L = index->getCandidatesEdgesNum();
    p.K = index->getResultEdgesNum();
    p.R = P.get<unsigned>("R");
    p.iters = P.get<unsigned>("ITER");
    p.S = index->getInitEdgesNum() && !efanna_ ? index->getInitEdgesNum() : P.get<unsigned>("S");
    p.dtype = dtype_;
    p.rerank_mult = P.has("rerank_mult") ? P.get<unsigned>("rerank_mult") : 0;
    p.cand_slots = P.has("cand_slots") ? P.get<unsigned>("cand_slots") : 0;
    p.seed = index->seed();
    const unsigned N = index->getBaseLen();
    auto &fg = index->getFinalGraph();
    std::vector<uint32_t> init, cnt;
    unsigned stride = p.S;
    if (fg.size() == N) {  // initial lists from the init stage (component_refine.cpp:96-103, :1046-1056)
        if (efanna_) {
            stride = 1;
            for (const auto &r : fg) stride = std::max<unsigned>(stride, (unsigned)r.size());
            stride = std::min(stride, p.L);
        }
        init.assign((size_t)N * stride, 0xFFFFFFFFu);
        cnt.assign(N, 0);
        for (unsigned i = 0; i < N; ++i) {
            std::vector<Index::SimpleNeighbor> row(fg[i].begin(), fg[i].end());
            if (efanna_ && row.size() > stride)
                std::partial_sort(row.begin(), row.begin() + stride, row.end(),
                                  [](const Index::SimpleNeighbor &a, const Index::SimpleNeighbor &b) {
                                      return a.distance < b.distance || (a.distance == b.distance && a.id < b.id);
                                  });
            const unsigned c = std::min<unsigned>(stride, (unsigned)row.size());
            for (unsigned j = 0; j < c; ++j) init[(size_t)i * stride + j] = row[j].id;
            cnt[i] = c;
        }
    }
    wv_index *ix = index->deviceIndex();
    if (efanna_) {
        index->check(wv_index_build_efanna(ix, &p, init.empty() ? nullptr : init.data(),
                                           cnt.empty() ? nullptr : cnt.data(), stride, &stats_));
    } else if (index->debug) {
        // NNDescent()'s debug path (component_refine.cpp:106-118): CONTROL_NUM
        // GenRandom control nodes, their exact top CONTROL_NUM including
        // themselves (generate_control_set, :241-256), the recall of each
        // round printed as eval_recall does (:258-275)
        const unsigned nc = std::min<unsigned>(CONTROL_NUM, N - 1);
        std::vector<unsigned> ctrl(nc);
        GenRandomStream(p.seed, TAG_CONTROL, 0, ctrl.data(), nc, N);
        const unsigned tk = std::min<unsigned>(CONTROL_NUM, N);
        std::vector<float> rows((size_t)nc * index->getBaseDim());
        for (unsigned i = 0; i < nc; ++i)
            std::copy(index->getBaseData() + (size_t)ctrl[i] * index->getBaseDim(),
                      index->getBaseData() + (size_t)(ctrl[i] + 1) * index->getBaseDim(),
                      rows.begin() + (size_t)i * index->getBaseDim());
        std::vector<uint32_t> truth((size_t)nc * tk);
        std::vector<double> td((size_t)nc * tk), per(p.iters);
        index->check(wv_exact_knn(index->ctx(), index->getBaseData(), N, rows.data(), nc, index->getBaseDim(), tk,
                                  -1, truth.data(), td.data(), nullptr));
        index->check(wv_index_build_nndescent_monitor(ix, &p, init.empty() ? nullptr : init.data(),
                                                      cnt.empty() ? nullptr : cnt.data(), ctrl.data(), nc,
                                                      truth.data(), tk, per.data(), &stats_));
        for (unsigned it = 0; it < p.iters; ++it) {
            std::cout << "NN-Descent iter: " << it << std::endl;
            std::cout << "Graph Quality : " << per[it] << std::endl;
        }
    } else {
        index->check(wv_index_build_nndescent(ix, &p, init.empty() ? nullptr : init.data(),
                                              cnt.empty() ? nullptr : cnt.data(), &stats_));
    }
    std::vector<uint32_t> ids((size_t)N * p.K), deg(N);
    std::vector<float> dist((size_t)N * p.K);
    index->check(wv_index_get_graph(ix, ids.data(), dist.data(), deg.data(), nullptr));
    fg.assign(N, {});
    for (unsigned i = 0; i < N; ++i) {
        fg[i].reserve(deg[i]);
        for (unsigned j = 0; j < deg[i]; ++j) fg[i].emplace_back(ids[(size_t)i * p.K + j], dist[(size_t)i * p.K + j]);
    }
    index->markLoadGraphChanged();  // the device graph is now the built one
}

// Upload FinalGraph as the device graph.
static void upload_final_graph(Index *index) {
    const unsigned N = index->getBaseLen();
    auto &fg = index->getFinalGraph();
    unsigned w = 1;
    for (const auto &r : fg) w = std::max<unsigned>(w, (unsigned)r.size());
    std::vector<uint32_t> ids((size_t)N * w, 0xFFFFFFFFu), deg(N);
    for (unsigned i = 0; i < N; ++i) {
        deg[i] = (uint32_t)fg[i].size();
        for (unsigned j = 0; j < deg[i]; ++j) ids[(size_t)i * w + j] = fg[i][j].id;
    }
    index->check(wv_index_set_graph(index->deviceIndex(), ids.data(), deg.data(), w));
}

// Device graph (W ids a row) back into FinalGraph; search takes it next.
static void download_final_graph(Index *index, unsigned W) {
    const unsigned N = index->getBaseLen();
    std::vector<uint32_t> gid((size_t)N * W), gdeg(N);
    std::vector<float> gd((size_t)N * W);
    index->check(wv_index_get_graph(index->deviceIndex(), gid.data(), gd.data(), gdeg.data(), nullptr));
    auto &fg = index->getFinalGraph();
    fg.assign(N, {});
    for (unsigned i = 0; i < N; ++i) {
        fg[i].reserve(gdeg[i]);
        for (unsigned j = 0; j < gdeg[i]; ++j) fg[i].emplace_back(gid[(size_t)i * W + j], gd[(size_t)i * W + j]);
    }
    index->getLoadGraph().clear();
    index->markLoadGraphChanged();
}

void ComponentRefineNSGHIP::RefineInner() {
    auto &P = index->getParam();
    wv_nsg_params p{};
    p.R = P.get<unsigned>("R_refine");  // SetConfigs, component_refine.cpp:430-436
    p.L = P.get<unsigned>("L_refine");
    p.C = P.get<unsigned>("C_refine");
    p.seed = index->seed();
    if (!index->ep_set_) {
        std::cout << "__ENTRY : Centroid__" << std::endl;
        ComponentRefineEntryCentroidHIP(index).EntryInner();
    }
    p.ep = index->ep_;
    if (index->getFinalGraph().size() != index->getBaseLen())
        throw std::invalid_argument("NSG refines a kNN graph: run init + refine first");
    upload_final_graph(index);
    std::cout << "__CANDIDATE : GREEDY(NSG)__ __PRUNE : NSG__ __CONN : DFS__" << std::endl;
    index->check(wv_index_build_nsg(index->deviceIndex(), &p, &stats_));
    download_final_graph(index, stats_.width);
}

void ComponentRefineSSGHIP::RefineInner() {
    auto &P = index->getParam();
    wv_ssg_params p{};
    p.R = P.get<unsigned>("R_refine");
    p.L = P.get<unsigned>("L_refine");
    p.angle = P.get<float>("A");
    p.n_try = P.get<unsigned>("n_try");
    p.seed = index->seed();
    if (index->getFinalGraph().size() != index->getBaseLen())
        throw std::invalid_argument("SSG refines a kNN graph: run init + refine first");
    upload_final_graph(index);
    std::cout << "__CANDIDATE : PROPAGATION 2__ __PRUNE : NSSG__ __CONN : DFS_Expand__" << std::endl;
    index->eps_.assign(p.n_try, 0);
    index->check(wv_index_build_ssg(index->deviceIndex(), &p, index->eps_.data(), &stats_));
    download_final_graph(index, stats_.width);
}

void ComponentRefineDPGHIP::RefineInner() {
    const unsigned range = index->getParam().get<unsigned>("K") / 2;  // SetConfigs, component_refine.cpp:756-764
    if (index->getFinalGraph().size() != index->getBaseLen())
        throw std::invalid_argument("DPG refines a kNN graph: run init + refine first");
    upload_final_graph(index);
    std::cout << "__PRUNE : DPG__ __CONN : REVERSE__" << std::endl;
    index->check(wv_index_build_dpg(index->deviceIndex(), range, &stats_));
    download_final_graph(index, stats_.width);
}

void ComponentRefineVamanaHIP::RefineInner() {
    auto &P = index->getParam();
    wv_vamana_params p{};
    p.R = P.get<unsigned>("R_refine");  // SetConfigs, component_refine.cpp:831-833
    p.L = P.get<unsigned>("L_refine");
    p.C = P.has("C_refine") ? P.get<unsigned>("C_refine") : 1024;
    p.alpha = index->alpha;
    p.seed = index->seed();
    if (!index->ep_set_) {
        std::cout << "__ENTRY : Centroid__" << std::endl;
        ComponentRefineEntryCentroidHIP(index).EntryInner();
    }
    p.ep = index->ep_;
    if (index->getFinalGraph().size() != index->getBaseLen())
        throw std::invalid_argument("Vamana refines a graph: run an init stage first");
    upload_final_graph(index);
    std::cout << "alpha " << p.alpha << "\n__CANDIDATE : NSG__ __PRUNE : VAMANA__" << std::endl;
    index->check(wv_index_build_vamana(index->deviceIndex(), &p, &stats_));
    download_final_graph(index, stats_.width);
    index->alpha = 2;  // the next pass (component_refine.cpp:871)
}

// ------------------------------------------------------------------ search
void ComponentSearchEntryRand::SearchEntryInner(unsigned query, std::vector<Index::Neighbor> &pool) {
    const unsigned L = index->getParam().get<unsigned>("L_search");
    pool.resize(L + 1);
    std::vector<unsigned> ids(L);
    GenRandomStream(index->seed(), TAG_ENTRY, query, ids.data(), L, index->getBaseLen());
    for (unsigned i = 0; i < L; ++i) pool[i] = Index::Neighbor(ids[i], 0.0f, true);
}

void ComponentRefineEntryCentroidHIP::EntryInner() {
    const unsigned N = index->getBaseLen(), d = index->getBaseDim();
    std::vector<double> sum(d, 0.0);
    const float *b = index->getBaseData();
    for (size_t i = 0; i < N; ++i)
        for (unsigned j = 0; j < d; ++j) sum[j] += b[i * d + j];
    std::vector<float> center(d);
    for (unsigned j = 0; j < d; ++j) center[j] = (float)(sum[j] / N);
    uint32_t id = 0;
    double dist = 0.0;
    index->check(wv_exact_knn(index->ctx(), b, N, center.data(), 1, d, 1, -1, &id, &dist, nullptr));
    index->ep_ = id;
    index->ep_set_ = true;
    std::cout << "ep_ " << index->ep_ << std::endl;
}

void ComponentSearchEntryCentroid::SearchEntryInner(unsigned query, std::vector<Index::Neighbor> &pool) {
    const unsigned L = index->getParam().get<unsigned>("L_search");
    if (!index->ep_set_)
        throw std::invalid_argument("no entry point: load an NSG/Vamana graph or run ComponentRefineEntryCentroidHIP");
    const auto &lg = index->getLoadGraph();
    static const std::vector<unsigned> none;
    head_then_fill(index->ep_ < lg.size() ? lg[index->ep_] : none, stream_of(index->seed(), TAG_CENTROID, query), L,
                   index->getBaseLen(), pool);
}

void ComponentSearchEntrySubCentroid::SearchEntryInner(unsigned query, std::vector<Index::Neighbor> &pool) {
    const unsigned L = index->getParam().get<unsigned>("L_search");
    if (index->eps_.size() > L) throw std::invalid_argument("more entry points (eps_) than L_search");
    head_then_fill(index->eps_, stream_of(index->seed(), TAG_SUBCENTROID, query), L, index->getBaseLen(), pool);
}

void ComponentSearchRoute::RouteBatch(unsigned nq, const std::vector<std::vector<Index::Neighbor>> *pools,
                                      std::vector<std::vector<unsigned>> &res) {
    if (!pools) throw std::invalid_argument("RouteBatch needs entry pools");
    res.resize(nq);
    for (unsigned i = 0; i < nq; ++i) {
        auto pool = (*pools)[i];
        RouteInner(i, pool, res[i]);
    }
}

void ComponentSearchRouteGreedyHIP::RouteInner(unsigned query, std::vector<Index::Neighbor> &pool,
                                               std::vector<unsigned> &res) {
    const unsigned L = index->getParam().get<unsigned>("L_search");  // route.cpp:11-12
    const unsigned K = index->getParam().get<unsigned>("K_search");
    if (pool.size() < L) throw std::invalid_argument("entry pool shorter than L_search");
    index->uploadLoadGraph();
    std::vector<uint32_t> ent(L), ids(L), ev(1), hp(1);
    std::vector<float> dist(L);
    for (unsigned i = 0; i < L; ++i) ent[i] = pool[i].id;
    const float *q = index->getQueryData() + (size_t)query * index->getQueryDim();
    wv_search_stats st{};
    index->check(wv_index_search(index->deviceIndex(), q, 1, ent.data(), 0, L, K, nullptr, nullptr, L,
                                 ids.data(), dist.data(), ev.data(), hp.data(), &st));
    for (unsigned i = 0; i < L; ++i) pool[i] = Index::Neighbor(ids[i], dist[i], false);
    res.assign(ids.begin(), ids.begin() + K);
    index->addDistCount(ev[0]);
    index->addHopCount(hp[0]);
    kernel_ms_ = st.kernel_ms;
}

void ComponentSearchRouteGreedyHIP::RouteBatch(unsigned nq, const std::vector<std::vector<Index::Neighbor>> *pools,
                                               std::vector<std::vector<unsigned>> &res) {
    const unsigned L = index->getParam().get<unsigned>("L_search");
    const unsigned K = index->getParam().get<unsigned>("K_search");
    index->uploadLoadGraph();
    std::vector<uint32_t> ent;
    if (pools) {  // caller-supplied entries: one launch with explicit entry sets
        if (pools->size() < nq) throw std::invalid_argument("fewer entry pools than queries");
        ent.resize((size_t)nq * L);
        for (unsigned i = 0; i < nq; ++i) {
            if ((*pools)[i].size() < L) throw std::invalid_argument("entry pool shorter than L_search");
            for (unsigned j = 0; j < L; ++j) ent[(size_t)i * L + j] = (*pools)[i][j].id;
        }
    }
    std::vector<uint32_t> ids((size_t)nq * K), ev(nq), hp(nq);
    wv_search_stats st{};
    index->check(wv_index_search(index->deviceIndex(), index->getQueryData(), nq, pools ? ent.data() : nullptr,
                                 index->seed(), L, K, nullptr, nullptr, K, ids.data(), nullptr, ev.data(), hp.data(),
                                 &st));
    res.assign(nq, {});
    for (unsigned i = 0; i < nq; ++i) res[i].assign(ids.begin() + (size_t)i * K, ids.begin() + (size_t)(i + 1) * K);
    index->addDistCount(st.dist_evals);
    index->addHopCount(st.hops);
    kernel_ms_ = st.kernel_ms;
}

}  // namespace weavess
