/*
 * wv_oracle.cpp -- CPU restatement of WEAVESS's kNN-graph hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see wv_oracle.h).  Built with
 * -O3 -fopenmp -ffp-contract=off so single-precision arithmetic follows the
 * source order exactly.
 *
 * Function                      follows (reference path:line)                         pinned by (tests/)
 * ----------------------------  ----------------------------------------------------  -----------------------------------------------
 * wvo_dist_ref                  include/weavess/distance.h:12-36                      oracle/_ref build of distance.h (bit-exact),
 *                                                                                     closed form on integer data
 * wvo_dist_canon                build order (D4), reduction tree below                FP64 recompute <=1e-5 rel, integer data exact,
 *                                                                                     tiny hand-worked vectors
 * wvo_gen_random                include/weavess/util.h:13-27                          oracle/_ref build of util.h fed the same draws,
 *                                                                                     invariants (distinct, in range, count)
 * wvo_init_random               src/component_init.cpp:11-45                          invariants (no self, distinct, sorted, dist)
 * wvo_nndescent                 src/component_refine.cpp:12-239,                      a 6-point hand-worked case: pools, flags, M,
 *                               include/weavess/index.h:86-150,                       nn_new/nn_old and pair counts after 1-3 rounds,
 *                               src/component_prune.cpp:8-18                          and the reference's single-thread sweep (D5)
 *                                                                                     (test_oracle_nnd_kat.py); S>=N-1 gives the
 *                                                                                     exact kNN graph; invariants; graph recall
 *                                                                                     vs FP64 exact kNN
 * wvo_nndescent_state           the same rounds, state before the final sort          the hand-worked case above
 * wvo_nndescent_efanna          src/component_refine.cpp:959-1183,                    ITER=0 returns the initial rows sorted, self
 *                               include/weavess/index.h:105-110                       and repeats dropped (numpy); 2S=N-1 with exact
 *                                                                                     insertion gives the exact kNN graph in one
 *                                                                                     round (the samples are joined); graph recall
 * wvo_nsg                       src/component_refine.cpp:393-540,                     every node a candidate gives the MRNG selection
 *                               component_candidate.cpp:10-85,                        (numpy brute force) plus reverse edges; a degree
 *                               component_prune.cpp:20-66,                            cap re-prunes as stated; every node reachable
 *                               component_conn.cpp:10-140
 * wvo_nsg_conn                  component_conn.cpp:14-98                              two disjoint cliques get exactly one link, from the
 *                                                                                     nearest reached node
 * wvo_vamana                    src/component_refine.cpp:794-958,                     huge alpha gives the exact R-NN graph; every node
 *                               component_prune.cpp:245-297                           a candidate gives numpy RobustPrune + fill
 * wvo_ssg                       src/component_refine.cpp:550-727,                     threshold = cos(A); a hand-built star keeps rays
 *                               component_candidate.cpp:89-108,                       60 degrees apart; two cliques get the links the
 *                               component_prune.cpp:71-122,                           rule names; every node reachable from every root
 *                               component_conn.cpp:169-226
 * wvo_dpg                       src/component_refine.cpp:728-790,                     points on a line (closed-form occlusion counts);
 *                               component_prune.cpp:124-186,                          rows = numpy selection + reverse edges, sorted
 *                               component_conn.cpp:228-268
 * wvo_search_greedy             src/component_search_entry.cpp:9-30,                  complete graph / all-node entry set / L>=N give
 *                               src/component_route.cpp:9-53,                         exact kNN; hand-worked 8-node KAT for the
 *                               include/weavess/index.h:152-180                       k-pointer restart and duplicate rejection
 * wvo_exact_knn                 algo/knn/gen_knn.cpp:117-150 (self skipped properly)  numpy argsort brute force on small inputs
 * wvo_recall                    src/builder.cpp:465-480                               hand-computed examples
 * wvo_graph_quality             src/builder.cpp:760-776                               hand-computed examples
 *
 * Learned early termination has no reference counterpart (ml/README.md only
 * links papers); these restate the device's formulas and are pinned as follows:
 * wvo_term_features             weavess_hip.h wv_term_model feature list              hand-worked vector; zero-distance clamps
 * wvo_search_greedy_ext         wvo_search_greedy + features / settle / stop          plain-Python sequential search on integer
 *                                                                                     data (features, settle, results); always-stop
 *                                                                                     model stops after one expansion
 * wvo_mlp_grad                  FP64 backprop of weighted BCE on the logit            closed form (zero weights); central finite
 *                                                                                     differences of its own loss
 * wvo_mlp_logit                 FP64 forward; BF16-product order of the device        closed form (one ReLU unit, BF16-exact
 *                                                                                     inputs); BF16 within 5 % of FP64
 */
#include "wv_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>
#include <omp.h>

namespace {

/* ---------------------------------------------------------------------- */
/* Keys: (dist, id) packed so that unsigned order == (dist, id) order for  */
/* non-negative finite dist.                                               */
/* ---------------------------------------------------------------------- */
inline uint64_t make_key(float d, uint32_t id) {
    uint32_t b;
    std::memcpy(&b, &d, 4);
    return ((uint64_t)b << 32) | id;
}
inline uint32_t key_id(uint64_t k) { return (uint32_t)k; }
inline float key_dist(uint64_t k) {
    uint32_t b = (uint32_t)(k >> 32);
    float d;
    std::memcpy(&d, &b, 4);
    return d;
}
constexpr uint64_t KEY_EMPTY = ~0ull;

/* splitmix64 finaliser (Steele, Lea, Flood 2014). */
inline uint64_t fmix64(uint64_t z) {
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    return z;
}
constexpr uint64_t GOLDEN = 0x9E3779B97F4A7C15ull;

/* murmur3 fmix32. */
inline uint32_t fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

/* Stream tags (shared with the GPU path by value). */
constexpr uint64_t TAG_INIT = 0x494E4954ull;    /* "INIT" */
constexpr uint64_t TAG_ENTRY = 0x454E5452ull;   /* "ENTR" */
constexpr uint64_t TAG_REVNEW = 0x52564E57ull;  /* "RVNW" */
constexpr uint64_t TAG_REVOLD = 0x52564F4Cull;  /* "RVOL" */
constexpr uint64_t TAG_EFANNA = 0x45464E41ull;  /* "EFNA": first-round samples */

inline uint64_t stream_of(uint64_t seed, uint64_t tag, uint64_t idx) {
    uint64_t h = fmix64(seed + GOLDEN);
    h = fmix64(h ^ (tag * GOLDEN));
    h = fmix64(h ^ (idx + 0x632BE59BD9B4E019ull));
    return h;
}
inline uint32_t draw(uint64_t stream, uint64_t k) {
    return (uint32_t)(fmix64(stream + (k + 1) * GOLDEN) >> 32);
}

/* Slot of candidate `id` in target `t`'s table (GPU: same formula). */
inline uint32_t cand_slot(uint32_t t, uint32_t id, uint32_t mask) {
    return fmix32(id * 0x9E3779B1u + t * 0x7F4A7C15u) & mask;
}

/* Build-order distance (D4).  Chunks of 4 consecutive elements; chunk c
 * contributes s = ((x0^2 + x1^2) + x2^2) + x3^2 (the grouping of
 * distance.h:24) to accumulator c % 16; accumulators are then folded
 * acc[j] += acc[j + h] for h = 8, 4, 2, 1.  A trailing partial chunk is
 * zero-padded, which adds exact zeros.  A GPU tile of 16/m lanes per row in
 * which lane j owns accumulators j, j + 16/m, ... reproduces this exactly. */
inline float dist_canon(const float *a, const float *b, uint32_t d) {
    float acc[16];
    for (int j = 0; j < 16; ++j) acc[j] = 0.0f;
    uint32_t nch = d / 4;
    for (uint32_t c = 0; c < nch; ++c) {
        const float *x = a + 4 * c, *y = b + 4 * c;
        float d0 = x[0] - y[0], d1 = x[1] - y[1], d2 = x[2] - y[2], d3 = x[3] - y[3];
        float s = d0 * d0 + d1 * d1;
        s = s + d2 * d2;
        s = s + d3 * d3;
        acc[c & 15] = acc[c & 15] + s;
    }
    uint32_t rem = d & 3;
    if (rem) {
        const float *x = a + 4 * nch, *y = b + 4 * nch;
        float s = 0.0f;
        float d0 = x[0] - y[0];
        s = d0 * d0;
        if (rem > 1) { float d1 = x[1] - y[1]; s = s + d1 * d1; }
        if (rem > 2) { float d2 = x[2] - y[2]; s = s + d2 * d2; }
        acc[nch & 15] = acc[nch & 15] + s;
    }
    for (int h = 8; h >= 1; h >>= 1)
        for (int j = 0; j < h; ++j) acc[j] = acc[j] + acc[j + h];
    return acc[0];
}

int nthreads_or_default(int n) { return n > 0 ? n : omp_get_max_threads(); }

/* GenRandom (include/weavess/util.h:13-27): `size` draws reduced mod
 * (N - size), sorted, forced strictly increasing, then rotated by one more
 * draw mod N.  Draw i is next(i). */
template <class Next>
void gen_random_with(Next next, uint32_t *addr, uint32_t size, uint32_t N) {
    for (uint32_t i = 0; i < size; ++i) addr[i] = next(i) % (N - size);
    std::sort(addr, addr + size);
    for (uint32_t i = 1; i < size; ++i)
        if (addr[i] <= addr[i - 1]) addr[i] = addr[i - 1] + 1;
    uint32_t off = next(size) % N;
    for (uint32_t i = 0; i < size; ++i) addr[i] = (uint32_t)(((uint64_t)addr[i] + off) % N);
}
/* ... on the counter stream (D2). */
void gen_random(uint64_t stream, uint32_t *addr, uint32_t size, uint32_t N) {
    gen_random_with([stream](uint32_t i) { return draw(stream, i); }, addr, size, N);
}

}  // namespace

extern "C" {

float wvo_dist_ref(const float *a, const float *b, uint32_t d) {
    /* distance.h:16-35: groups of four while at least four remain, each
     * group summed left to right and then added to the running total; the
     * remaining 0-3 elements are added one at a time. */
    float result = 0.0f;
    uint32_t i = 0;
    for (; i + 4 <= d; i += 4) {
        float d0 = a[i] - b[i], d1 = a[i + 1] - b[i + 1];
        float d2 = a[i + 2] - b[i + 2], d3 = a[i + 3] - b[i + 3];
        float g = d0 * d0 + d1 * d1;
        g = g + d2 * d2;
        g = g + d3 * d3;
        result = result + g;
    }
    for (; i < d; ++i) {
        float t = a[i] - b[i];
        result = result + t * t;
    }
    return result;
}

float wvo_dist_canon(const float *a, const float *b, uint32_t d) { return dist_canon(a, b, d); }

double wvo_dist_f64(const float *a, const float *b, uint32_t d) {
    double s = 0.0;
    for (uint32_t i = 0; i < d; ++i) {
        double t = (double)a[i] - (double)b[i];
        s += t * t;
    }
    return s;
}

uint64_t wvo_stream(uint64_t seed, uint64_t tag, uint64_t idx) { return stream_of(seed, tag, idx); }
uint32_t wvo_draw(uint64_t stream, uint64_t k) { return draw(stream, k); }

int wvo_gen_random(uint64_t stream, uint32_t *addr, uint32_t size, uint32_t N) {
    if (!addr || size >= N) return -1;
    gen_random(stream, addr, size, N);
    return 0;
}

int wvo_gen_random_draws(const uint32_t *draws, uint32_t *addr, uint32_t size, uint32_t N) {
    if (!draws || !addr || size >= N) return -1;
    gen_random_with([draws](uint32_t i) { return draws[i]; }, addr, size, N);
    return 0;
}

int wvo_init_random(const float *base, uint32_t N, uint32_t d, uint32_t S, uint64_t seed,
                    uint32_t *ids, float *dist, uint32_t *cnt, int nthreads) {
    /* src/component_init.cpp:11-45: per node, GenRandom(S) over [0,N), skip
     * self, FP32 distances, ascending sort (here by (dist,id), D1). */
    if (!base || !ids || !dist || !cnt || S == 0 || S >= N) return -1;
#pragma omp parallel for num_threads(nthreads_or_default(nthreads)) schedule(static)
    for (int64_t ii = 0; ii < (int64_t)N; ++ii) {
        uint32_t i = (uint32_t)ii;
        std::vector<uint32_t> tmp(S);
        gen_random(stream_of(seed, TAG_INIT, i), tmp.data(), S, N);
        std::vector<uint64_t> keys;
        keys.reserve(S);
        for (uint32_t j = 0; j < S; ++j) {
            if (tmp[j] == i) continue;
            float dd = dist_canon(base + (size_t)i * d, base + (size_t)tmp[j] * d, d);
            keys.push_back(make_key(dd, tmp[j]));
        }
        std::sort(keys.begin(), keys.end());
        cnt[i] = (uint32_t)keys.size();
        for (uint32_t j = 0; j < S; ++j) {
            if (j < keys.size()) {
                ids[(size_t)i * S + j] = key_id(keys[j]);
                dist[(size_t)i * S + j] = key_dist(keys[j]);
            } else {
                ids[(size_t)i * S + j] = 0xFFFFFFFFu;
                dist[(size_t)i * S + j] = -1.0f;
            }
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* NN-Descent                                                              */
/* ---------------------------------------------------------------------- */
namespace {

struct Ent {
    uint64_t key;
    uint8_t flag;
};
struct EntLess {
    bool operator()(const Ent &a, const Ent &b) const { return a.key < b.key; }
};

struct NndState {
    uint32_t N, d, L, S, R;
    bool sweep = false;
    const float *base;
    std::vector<std::vector<Ent>> pool;      /* max-heap during join, sorted otherwise */
    std::vector<std::mutex> locks;
    std::vector<uint32_t> M;
    std::vector<std::vector<uint32_t>> nn_new, nn_old;
    /* slot mode */
    uint32_t C = 0;
    std::vector<uint64_t> slots;             /* N*C */
    std::vector<uint64_t> thr;               /* per-node threshold key */
    std::vector<std::vector<uint32_t>> pool_ids;  /* sorted pool ids at join start */
    NndState(uint32_t n) : pool(n), locks(n), M(n), nn_new(n), nn_old(n) {}
};

/* nhood::insert, include/weavess/index.h:120-135, with (dist,id) order (D1):
 * reject if worse than the current worst; reject a duplicate id; push while
 * below capacity L, else replace the worst. */
bool insert_exact(NndState &st, uint32_t t, uint32_t id, float dd) {
    uint64_t key = make_key(dd, id);
    std::lock_guard<std::mutex> g(st.locks[t]);
    auto &p = st.pool[t];
    if (!p.empty() && key > p.front().key) return false;
    for (const Ent &e : p)
        if (key_id(e.key) == id) return false;
    if (p.size() < st.L) {
        p.push_back(Ent{key, 1});
        std::push_heap(p.begin(), p.end(), EntLess());
    } else {
        std::pop_heap(p.begin(), p.end(), EntLess());
        p.back() = Ent{key, 1};
        std::push_heap(p.begin(), p.end(), EntLess());
    }
    return true;
}

/* Lossy slot-table insertion mirrored by the GPU: keep the smaller key per
 * slot; the threshold is the target's worst key at the start of the join. */
bool insert_slot(NndState &st, uint32_t t, uint32_t id, float dd) {
    uint64_t key = make_key(dd, id);
    if (key > st.thr[t]) return false;
    /* an id already in the target's pool is rejected before it can take a
     * slot (the duplicate test of index.h:123-125, done up front) */
    const auto &ids = st.pool_ids[t];
    if (std::binary_search(ids.begin(), ids.end(), id)) return false;
    uint64_t *slot = &st.slots[(size_t)t * st.C + cand_slot(t, id, st.C - 1)];
    uint64_t cur = __atomic_load_n(slot, __ATOMIC_RELAXED);
    while (key < cur && !__atomic_compare_exchange_n(slot, &cur, key, true, __ATOMIC_RELAXED,
                                                      __ATOMIC_RELAXED)) {
    }
    return true;
}

/* join(), src/component_refine.cpp:127-143 with nhood::join,
 * include/weavess/index.h:137-149. */
void nnd_join(NndState &st, int nth, uint64_t &pairs, uint64_t &accepted) {
    const bool slots = st.C > 0;
    if (slots) {
#pragma omp parallel for num_threads(nth) schedule(static)
        for (int64_t t = 0; t < (int64_t)st.N; ++t) {
            const auto &p = st.pool[t];
            st.thr[t] = p.empty() ? KEY_EMPTY : p.back().key;  /* pools are sorted here */
            auto &ids = st.pool_ids[t];
            ids.clear();
            for (const Ent &e : p) ids.push_back(key_id(e.key));
            std::sort(ids.begin(), ids.end());
        }
    } else {
#pragma omp parallel for num_threads(nth) schedule(static)
        for (int64_t t = 0; t < (int64_t)st.N; ++t)
            std::make_heap(st.pool[t].begin(), st.pool[t].end(), EntLess());
    }
    uint64_t pc = 0, ac = 0;
#pragma omp parallel for num_threads(nth) schedule(dynamic, 100) reduction(+ : pc, ac)
    for (int64_t nn = 0; nn < (int64_t)st.N; ++nn) {
        const auto &nw = st.nn_new[nn];
        const auto &od = st.nn_old[nn];
        auto cb = [&](uint32_t i, uint32_t j) {
            if (i == j) return;
            float dd = dist_canon(st.base + (size_t)i * st.d, st.base + (size_t)j * st.d, st.d);
            ++pc;
            if (slots) {
                ac += insert_slot(st, i, j, dd);
                ac += insert_slot(st, j, i, dd);
            } else {
                ac += insert_exact(st, i, j, dd);
                ac += insert_exact(st, j, i, dd);
            }
        };
        for (uint32_t i : nw) {
            for (uint32_t j : nw)
                if (i < j) cb(i, j);
            for (uint32_t j : od) cb(i, j);
        }
    }
    pairs = pc;
    accepted = ac;
    if (slots) {
        /* merge: union of the pool and the slot winners, duplicates by id
         * dropped (the pool entry wins), ascending, first L kept; winners are
         * flagged new.  Slots are reset for the next join. */
#pragma omp parallel for num_threads(nth) schedule(dynamic, 256)
        for (int64_t tt = 0; tt < (int64_t)st.N; ++tt) {
            auto &p = st.pool[tt];
            uint64_t *sl = &st.slots[(size_t)tt * st.C];
            std::vector<Ent> u(p.begin(), p.end());
            for (uint32_t s = 0; s < st.C; ++s) {
                uint64_t k = sl[s];
                sl[s] = KEY_EMPTY;
                if (k == KEY_EMPTY) continue;
                bool dup = false;
                for (const Ent &e : p)
                    if (key_id(e.key) == key_id(k)) { dup = true; break; }
                if (!dup) u.push_back(Ent{k, 1});
            }
            std::sort(u.begin(), u.end(), EntLess());
            if (u.size() > st.L) u.resize(st.L);
            p.swap(u);
        }
    }
}

/* Reverse-sample priority key of proposer n at node x: a 32-bit hash in the
 * high word, n in the low word; the R smallest keys are kept (D3). */
uint64_t rev_key(uint64_t seed, uint64_t tag, uint32_t iter, uint32_t x, uint32_t n) {
    uint64_t h = fmix64(stream_of(seed, tag, ((uint64_t)iter << 32) | x) ^ ((uint64_t)n * GOLDEN));
    return ((h >> 32) << 32) | n;
}

/* update(), src/component_refine.cpp:145-239, two-phase (all pools sorted
 * before any reverse test) and with the reverse cap of D3. */
void nnd_update(NndState &st, int nth, uint32_t iter, uint64_t seed) {
    const uint32_t N = st.N, S = st.S, R = st.R;
    std::vector<float> worst(N), leaf(N);
    const bool sweep = st.sweep;
    /* :157-175 sort, truncate to L, new M */
#pragma omp parallel for num_threads(nth) schedule(static)
    for (int64_t n = 0; n < (int64_t)N; ++n) {
        auto &p = st.pool[n];
        std::sort(p.begin(), p.end(), EntLess());
        if (p.size() > st.L) p.resize(st.L);
        uint32_t maxl = std::min(st.M[n] + S, (uint32_t)p.size());
        uint32_t c = 0, l = 0;
        while (l < maxl && c < S) {
            if (p[l].flag) ++c;
            ++l;
        }
        st.M[n] = l;
        worst[n] = p.empty() ? -1.0f : key_dist(p.back().key);
        if (sweep) {
            /* what pool.back() reads once n's own sampling re-heaped it
             * (src/component_refine.cpp:212) */
            std::vector<Ent> h(p.begin(), p.end());
            std::make_heap(h.begin(), h.end(), EntLess());
            leaf[n] = h.empty() ? -1.0f : key_dist(h.back().key);
        }
        st.nn_new[n].clear();
        st.nn_old[n].clear();
    }
    /* :176-209 forward samples and reverse proposals */
    std::vector<std::atomic<uint32_t>> cnt_new(N), cnt_old(N);
    for (uint32_t n = 0; n < N; ++n) { cnt_new[n] = 0; cnt_old[n] = 0; }
#pragma omp parallel for num_threads(nth) schedule(static)
    for (int64_t n = 0; n < (int64_t)N; ++n) {
        auto &p = st.pool[n];
        for (uint32_t l = 0; l < st.M[n]; ++l) {
            uint32_t id = key_id(p[l].key);
            bool rev = key_dist(p[l].key) > ((sweep && id < (uint32_t)n) ? leaf[id] : worst[id]);
            if (p[l].flag) {
                st.nn_new[n].push_back(id);
                if (rev) cnt_new[id].fetch_add(1, std::memory_order_relaxed);
            } else {
                st.nn_old[n].push_back(id);
                if (rev) cnt_old[id].fetch_add(1, std::memory_order_relaxed);
            }
        }
    }
    std::vector<uint64_t> off_new(N + 1, 0), off_old(N + 1, 0);
    for (uint32_t n = 0; n < N; ++n) {
        off_new[n + 1] = off_new[n] + cnt_new[n].load();
        off_old[n + 1] = off_old[n] + cnt_old[n].load();
        cnt_new[n] = 0;
        cnt_old[n] = 0;
    }
    std::vector<uint32_t> rv_new(off_new[N]), rv_old(off_old[N]);
#pragma omp parallel for num_threads(nth) schedule(static)
    for (int64_t n = 0; n < (int64_t)N; ++n) {
        auto &p = st.pool[n];
        for (uint32_t l = 0; l < st.M[n]; ++l) {
            uint32_t id = key_id(p[l].key);
            if (!(key_dist(p[l].key) > ((sweep && id < (uint32_t)n) ? leaf[id] : worst[id]))) continue;
            if (p[l].flag)
                rv_new[off_new[id] + cnt_new[id].fetch_add(1, std::memory_order_relaxed)] = (uint32_t)n;
            else
                rv_old[off_old[id] + cnt_old[id].fetch_add(1, std::memory_order_relaxed)] = (uint32_t)n;
        }
        for (uint32_t l = 0; l < st.M[n]; ++l) p[l].flag = 0;  /* :189 nn.flag = false */
    }
    /* :213-238 append the (capped) reverse samples; cap old at 2R */
#pragma omp parallel for num_threads(nth) schedule(dynamic, 256)
    for (int64_t x = 0; x < (int64_t)N; ++x) {
        for (int which = 0; which < 2; ++which) {
            uint32_t *b = which == 0 ? &rv_new[off_new[x]] : &rv_old[off_old[x]];
            uint64_t c = which == 0 ? off_new[x + 1] - off_new[x] : off_old[x + 1] - off_old[x];
            std::vector<uint32_t> sel(b, b + c);
            if (c > R) {
                uint64_t tag = which == 0 ? TAG_REVNEW : TAG_REVOLD;
                std::vector<uint64_t> pr(c);
                for (uint64_t q = 0; q < c; ++q) pr[q] = rev_key(seed, tag, iter, (uint32_t)x, b[q]);
                std::nth_element(pr.begin(), pr.begin() + R, pr.end());
                sel.resize(R);
                for (uint32_t q = 0; q < R; ++q) sel[q] = (uint32_t)pr[q];
            }
            std::sort(sel.begin(), sel.end());
            auto &dst = which == 0 ? st.nn_new[x] : st.nn_old[x];
            dst.insert(dst.end(), sel.begin(), sel.end());
        }
        if (st.nn_old[x].size() > 2 * (size_t)R) st.nn_old[x].resize(2 * (size_t)R);
    }
}

/* ITER rounds of join() + update().  first_join: round 0 joins the new lists
 * already set (EFANNA's samples); else round 0 has nothing to join. */
void nnd_iterate(NndState &st, const wvo_nnd_params *p, int nth, bool first_join,
                 wvo_nnd_iter_stats *iter_stats) {
    for (uint32_t it = 0; it < p->iters; ++it) {
        uint64_t pairs = 0, acc = 0;
        if (it > 0 || first_join) nnd_join(st, nth, pairs, acc);
        nnd_update(st, nth, it, p->seed);
        if (iter_stats) {
            iter_stats[it].pairs = pairs;
            iter_stats[it].inserts = acc;
        }
    }
}

/* RefineInner's tail (:22-74) and PruneNaive: sort, keep the first K. */
void nnd_finalize(NndState &st, const wvo_nnd_params *p, int nth, uint32_t *out_ids, float *out_dist,
                  uint32_t *out_deg) {
#pragma omp parallel for num_threads(nth) schedule(static)
    for (int64_t i = 0; i < (int64_t)st.N; ++i) {
        auto &pl = st.pool[i];
        std::sort(pl.begin(), pl.end(), EntLess());
        uint32_t deg = std::min<uint32_t>(p->K, (uint32_t)pl.size());
        out_deg[i] = deg;
        for (uint32_t j = 0; j < p->K; ++j) {
            if (j < deg) {
                out_ids[(size_t)i * p->K + j] = key_id(pl[j].key);
                out_dist[(size_t)i * p->K + j] = key_dist(pl[j].key);
            } else {
                out_ids[(size_t)i * p->K + j] = 0xFFFFFFFFu;
                out_dist[(size_t)i * p->K + j] = -1.0f;
            }
        }
    }
}

/* KGraph's init() (src/component_refine.cpp:87-104): pools seeded from the
 * initial graph (or InitRandom's), all flagged new, M = S.  Returns false on
 * bad arguments. */
bool nnd_setup(NndState &st, const float *base, uint32_t N, uint32_t d, const wvo_nnd_params *p,
               const uint32_t *init_ids, const float *init_dist, const uint32_t *init_cnt, int nth) {
    if (!base || !p) return false;
    if (N < 2 || d == 0 || p->S == 0 || p->S >= N || p->L == 0 || p->K == 0 || p->K > p->L) return false;
    if (p->S > p->L) return false;
    if (p->cand_slots && (p->cand_slots & (p->cand_slots - 1))) return false;
    const uint32_t S = p->S;
    std::vector<uint32_t> ids0, cnt0;
    std::vector<float> dist0;
    if (!init_ids) {
        ids0.resize((size_t)N * S);
        dist0.resize((size_t)N * S);
        cnt0.resize(N);
        wvo_init_random(base, N, d, S, p->seed, ids0.data(), dist0.data(), cnt0.data(), nth);
        init_ids = ids0.data();
        init_dist = dist0.data();
        init_cnt = cnt0.data();
    }
    st.N = N; st.d = d; st.L = p->L; st.S = S; st.R = p->R; st.base = base;
    st.sweep = p->sweep != 0;
    st.C = p->cand_slots;
    if (st.C) {
        st.slots.assign((size_t)N * st.C, KEY_EMPTY);
        st.thr.assign(N, KEY_EMPTY);
        st.pool_ids.resize(N);
    }
    for (uint32_t i = 0; i < N; ++i) {
        auto &pl = st.pool[i];
        pl.reserve(p->L);
        uint32_t c = init_cnt ? init_cnt[i] : S;
        for (uint32_t j = 0; j < c && j < p->L; ++j)
            pl.push_back(Ent{make_key(init_dist[(size_t)i * S + j], init_ids[(size_t)i * S + j]), 1});
        std::sort(pl.begin(), pl.end(), EntLess());
        st.M[i] = S;
    }
    return true;
}

}  // namespace

int wvo_nndescent(const float *base, uint32_t N, uint32_t d, const wvo_nnd_params *p,
                  const uint32_t *init_ids, const float *init_dist, const uint32_t *init_cnt,
                  uint32_t *out_ids, float *out_dist, uint32_t *out_deg,
                  wvo_nnd_iter_stats *iter_stats, int nthreads) {
    if (!out_ids || !out_dist || !out_deg) return -1;
    const int nth = nthreads_or_default(nthreads);
    NndState st(N);
    if (!nnd_setup(st, base, N, d, p, init_ids, init_dist, init_cnt, nth)) return -1;
    /* NNDescent(), :106-125.  Iteration 0's join sees nn_new = 2S copies of
     * id 0 and no nn_old (index.h:98-103), so it evaluates nothing. */
    nnd_iterate(st, p, nth, false, iter_stats);
    nnd_finalize(st, p, nth, out_ids, out_dist, out_deg);
    return 0;
}

int wvo_nndescent_state(const float *base, uint32_t N, uint32_t d, const wvo_nnd_params *p,
                        const uint32_t *init_ids, const float *init_dist, const uint32_t *init_cnt,
                        uint32_t *pool_ids, float *pool_dist, uint8_t *pool_flag, uint32_t *pool_cnt,
                        uint32_t *M, uint32_t *nn_new, uint32_t *new_cnt, uint32_t new_cap,
                        uint32_t *nn_old, uint32_t *old_cnt, uint32_t old_cap,
                        wvo_nnd_iter_stats *iter_stats, int nthreads) {
    if (!pool_ids || !pool_dist || !pool_flag || !pool_cnt || !M || !nn_new || !new_cnt || !nn_old || !old_cnt)
        return -1;
    const int nth = nthreads_or_default(nthreads);
    NndState st(N);
    if (!nnd_setup(st, base, N, d, p, init_ids, init_dist, init_cnt, nth)) return -1;
    nnd_iterate(st, p, nth, false, iter_stats);
    for (uint32_t i = 0; i < N; ++i) {
        const auto &pl = st.pool[i];
        if (pl.size() > p->L || st.nn_new[i].size() > new_cap || st.nn_old[i].size() > old_cap) return -2;
        pool_cnt[i] = (uint32_t)pl.size();
        for (uint32_t j = 0; j < p->L; ++j) {
            const bool ok = j < pl.size();
            pool_ids[(size_t)i * p->L + j] = ok ? key_id(pl[j].key) : 0xFFFFFFFFu;
            pool_dist[(size_t)i * p->L + j] = ok ? key_dist(pl[j].key) : -1.0f;
            pool_flag[(size_t)i * p->L + j] = ok ? pl[j].flag : 0;
        }
        M[i] = st.M[i];
        new_cnt[i] = (uint32_t)st.nn_new[i].size();
        old_cnt[i] = (uint32_t)st.nn_old[i].size();
        for (uint32_t j = 0; j < new_cap; ++j)
            nn_new[(size_t)i * new_cap + j] = j < st.nn_new[i].size() ? st.nn_new[i][j] : 0xFFFFFFFFu;
        for (uint32_t j = 0; j < old_cap; ++j)
            nn_old[(size_t)i * old_cap + j] = j < st.nn_old[i].size() ? st.nn_old[i][j] : 0xFFFFFFFFu;
    }
    return 0;
}

int wvo_nndescent_efanna(const float *base, uint32_t N, uint32_t d, const wvo_nnd_params *p,
                         const uint32_t *init_ids, const uint32_t *init_cnt, uint32_t init_stride,
                         uint32_t *out_ids, float *out_dist, uint32_t *out_deg,
                         wvo_nnd_iter_stats *iter_stats, int nthreads) {
    if (!base || !p || !out_ids || !out_dist || !out_deg) return -1;
    if (N < 2 || d == 0 || p->S == 0 || 2 * (uint64_t)p->S >= N || p->L == 0 || p->K == 0 || p->K > p->L) return -1;
    if (p->cand_slots && (p->cand_slots & (p->cand_slots - 1))) return -1;
    if (init_ids && (init_stride == 0 || init_stride > p->L)) return -1;
    const int nth = nthreads_or_default(nthreads);
    const uint32_t S = p->S;
    std::vector<uint32_t> ids0, cnt0;
    std::vector<float> dist0;
    if (!init_ids) {  /* no initial graph given: InitRandom's */
        ids0.resize((size_t)N * S);
        dist0.resize((size_t)N * S);
        cnt0.resize(N);
        wvo_init_random(base, N, d, S, p->seed, ids0.data(), dist0.data(), cnt0.data(), nth);
        init_ids = ids0.data();
        init_cnt = cnt0.data();
        init_stride = S;
    }
    NndState st(N);
    st.N = N; st.d = d; st.L = p->L; st.S = S; st.R = p->R; st.base = base;
    st.C = p->cand_slots;
    if (st.C) {
        st.slots.assign((size_t)N * st.C, KEY_EMPTY);
        st.thr.assign(N, KEY_EMPTY);
        st.pool_ids.resize(N);
    }
    /* ComponentRefineEFANNA::init, src/component_refine.cpp:1036-1058: the
     * pool is the initial graph's row sorted, self and repeated ids skipped
     * (here every repeat, not only adjacent ones), all flagged new, M = S;
     * nn_new is GenRandom(2S) over [0, N) (index.h:105-110), nn_old empty. */
#pragma omp parallel for num_threads(nth) schedule(static)
    for (int64_t ii = 0; ii < (int64_t)N; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const uint32_t c = std::min(init_cnt ? init_cnt[i] : init_stride, init_stride);
        std::vector<uint64_t> keys;
        for (uint32_t j = 0; j < c; ++j) {
            const uint32_t id = init_ids[(size_t)i * init_stride + j];
            if (id >= N || id == i) continue;
            keys.push_back(make_key(dist_canon(base + (size_t)i * d, base + (size_t)id * d, d), id));
        }
        std::sort(keys.begin(), keys.end());
        keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
        if (keys.size() > p->L) keys.resize(p->L);
        auto &pl = st.pool[i];
        pl.clear();
        for (uint64_t k : keys) pl.push_back(Ent{k, 1});
        st.M[i] = S;
        st.nn_new[i].resize(2 * S);
        gen_random(stream_of(p->seed, TAG_EFANNA, i), st.nn_new[i].data(), 2 * S, N);
        st.nn_old[i].clear();
    }
    /* NNDescent(), :1059-1078: the first join runs over those samples. */
    nnd_iterate(st, p, nth, true, iter_stats);
    nnd_finalize(st, p, nth, out_ids, out_dist, out_deg);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Greedy search                                                           */
/* ---------------------------------------------------------------------- */
namespace {
inline float bf16r(float x) {
    /* round to nearest even on the upper 16 bits (finite inputs) */
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}
/* The device's BF16 inference step (wv_mlp.hpp term_logit). */
float logit_bf16(const wvo_term_model *m, const float *f) {
    float part[32];
    for (int j = 0; j < 32; ++j) {
        float acc = m->b1[j];
        for (int i = 0; i < 8; ++i) acc = acc + bf16r(m->w1[j][i]) * bf16r((f[i] - m->mean[i]) * m->inv_std[i]);
        const float h = acc > 0.0f ? acc : 0.0f;
        part[j] = bf16r(m->w2[j]) * bf16r(h);
    }
    for (int h = 16; h >= 1; h >>= 1)
        for (int j = 0; j < h; ++j) part[j] = part[j] + part[j + h];
    return part[0] + m->b2;
}
}  // namespace

void wvo_term_features(float hops, float kth, float best, float entry, float kth4, float stale, float L,
                       float k, float *f) {
    /* weavess_hip.h: g = [log2(1+hops), kth/entry, best/entry, log2(entry),
     * (kth4-kth)/entry, log2(L), log2(k), log2(1+stale)] */
    const float e = entry > 1e-30f ? entry : 1e-30f;
    f[0] = std::log2(1.0f + hops);
    f[1] = kth / e;
    f[2] = best / e;
    f[3] = std::log2(e);
    f[4] = (kth4 - kth) / e;
    f[5] = std::log2(L);
    f[6] = std::log2(k);
    f[7] = std::log2(1.0f + stale);
}

int wvo_mlp_logit(const wvo_term_model *m, const float *X, uint32_t n, int bf16, double *out) {
    if (!m || !X || !out) return -1;
    for (uint32_t s = 0; s < n; ++s) {
        const float *f = X + (size_t)s * 8;
        if (bf16) {
            out[s] = logit_bf16(m, f);
            continue;
        }
        double z = m->b2;
        for (int j = 0; j < 32; ++j) {
            double a = m->b1[j];
            for (int i = 0; i < 8; ++i) a += (double)m->w1[j][i] * (((double)f[i] - m->mean[i]) * m->inv_std[i]);
            z += (double)m->w2[j] * (a > 0 ? a : 0);
        }
        out[s] = z;
    }
    return 0;
}

int wvo_mlp_grad(const wvo_term_model *m, const float *X, const float *y, uint32_t n, float pos_weight,
                 double *grad, double *loss) {
    if (!m || !X || !y || !grad || n == 0) return -1;
    for (int p = 0; p < 321; ++p) grad[p] = 0.0;
    double L = 0.0;
    for (uint32_t s = 0; s < n; ++s) {
        double x[8], z1[32], h[32];
        for (int i = 0; i < 8; ++i) x[i] = ((double)X[(size_t)s * 8 + i] - m->mean[i]) * m->inv_std[i];
        double z = m->b2;
        for (int j = 0; j < 32; ++j) {
            double a = m->b1[j];
            for (int i = 0; i < 8; ++i) a += (double)m->w1[j][i] * x[i];
            z1[j] = a;
            h[j] = a > 0 ? a : 0;
            z += (double)m->w2[j] * h[j];
        }
        const double t = y[s], w = t > 0.5 ? pos_weight : 1.0;
        L += w * (std::max(z, 0.0) - z * t + std::log1p(std::exp(-std::fabs(z))));
        const double dz = w * (1.0 / (1.0 + std::exp(-z)) - t);
        for (int j = 0; j < 32; ++j) {
            const double dz1 = z1[j] > 0 ? dz * m->w2[j] : 0.0;
            for (int i = 0; i < 8; ++i) grad[j * 8 + i] += dz1 * x[i];
            grad[256 + j] += dz1;
            grad[288 + j] += dz * h[j];
        }
        grad[320] += dz;
    }
    for (int p = 0; p < 321; ++p) grad[p] /= n;
    if (loss) *loss = L / n;
    return 0;
}

int wvo_search_greedy(const uint32_t *graph, const uint32_t *deg, uint32_t deg_max,
                      const float *base, uint32_t N, uint32_t d, const float *queries, uint32_t Q,
                      const uint32_t *entries, uint64_t seed, uint32_t L, uint32_t k,
                      const uint32_t *L_per_q, const uint32_t *k_per_q, uint32_t k_stride,
                      uint32_t *out_ids, float *out_dist, uint32_t *evals, uint32_t *hops,
                      int nthreads) {
    return wvo_search_greedy_ext(graph, deg, deg_max, base, N, d, queries, Q, entries, seed, L, k,
                                 L_per_q, k_per_q, k_stride, out_ids, out_dist, evals, hops, nullptr,
                                 nullptr, nullptr, nullptr, 0, nthreads);
}

int wvo_search_greedy_ext(const uint32_t *graph, const uint32_t *deg, uint32_t deg_max,
                          const float *base, uint32_t N, uint32_t d, const float *queries,
                          uint32_t Q, const uint32_t *entries, uint64_t seed, uint32_t L,
                          uint32_t k, const uint32_t *L_per_q, const uint32_t *k_per_q,
                          uint32_t k_stride, uint32_t *out_ids, float *out_dist,
                          uint32_t *evals, uint32_t *hops, const wvo_term_model *model,
                          float *trace, uint32_t *trace_len, uint32_t *settle, uint32_t cap,
                          int nthreads) {
    if (!graph || !deg || !base || !queries || !out_ids || N == 0 || d == 0) return -1;
    /* validate all queries first: L >= k (src/builder.cpp:435-438), L < N
     * for random entries (GenRandom needs N - L > 0), L <= N otherwise */
    for (uint32_t q = 0; q < Q; ++q) {
        uint32_t Lq = L_per_q ? L_per_q[q] : L, kq = k_per_q ? k_per_q[q] : k;
        if (Lq == 0 || kq == 0 || kq > Lq || kq > k_stride) return -1;
        if (entries ? Lq > N : Lq >= N) return -1;
        if (entries && L_per_q) return -1;  /* explicit entries need a fixed L */
    }
    const int nth = nthreads_or_default(nthreads);
#pragma omp parallel num_threads(nth)
    {
        std::vector<uint32_t> stamp(N, 0);
        uint32_t epoch = 0;
        std::vector<uint64_t> key;
        std::vector<uint8_t> flag;
        std::vector<uint32_t> ent;
#pragma omp for schedule(dynamic, 16)
        for (int64_t qq = 0; qq < (int64_t)Q; ++qq) {
            const uint32_t q = (uint32_t)qq;
            const uint32_t Lq = L_per_q ? L_per_q[q] : L, kq = k_per_q ? k_per_q[q] : k;
            const float *qv = queries + (size_t)q * d;
            if (++epoch == 0) { std::fill(stamp.begin(), stamp.end(), 0); epoch = 1; }
            key.assign(Lq + 1, KEY_EMPTY);
            flag.assign(Lq + 1, 0);
            ent.resize(Lq);
            uint32_t ev = 0, hp = 0;
            /* SearchEntryRand: L random ids, distances, sort; entries are not
             * marked visited (component_search_entry.cpp:18-27). */
            if (entries) std::memcpy(ent.data(), entries + (size_t)q * Lq, Lq * 4);
            else gen_random(stream_of(seed, TAG_ENTRY, q), ent.data(), Lq, N);
            for (uint32_t i = 0; i < Lq; ++i) {
                key[i] = make_key(dist_canon(qv, base + (size_t)ent[i] * d, d), ent[i]);
                flag[i] = 1;
                ++ev;
            }
            std::sort(key.begin(), key.begin() + Lq);
            /* state features for traces / the learned stop (weavess_hip.h) */
            const bool watch = model || trace;
            const uint32_t kidx = (kq < Lq ? kq : Lq) - 1;
            float entry_d = key_dist(key[0]);
            float hist[4];  /* k-th distance 1..4 expansions ago */
            for (float &x : hist) x = key_dist(key[kidx]);
            uint32_t st = 0;
            /* RouteGreedy, component_route.cpp:16-47 */
            uint32_t kk = 0;
            while (kk < Lq) {
                uint32_t nk = Lq;
                bool expanded = false;
                if (flag[kk]) {
                    expanded = true;
                    flag[kk] = 0;
                    ++hp;
                    uint32_t u = key_id(key[kk]);
                    uint32_t du = deg[u] < deg_max ? deg[u] : deg_max;
                    for (uint32_t m = 0; m < du; ++m) {
                        uint32_t id = graph[(size_t)u * deg_max + m];
                        if (id >= N) continue;
                        if (stamp[id] == epoch) continue;
                        stamp[id] = epoch;
                        float dd = dist_canon(qv, base + (size_t)id * d, d);
                        ++ev;
                        uint64_t nkey = make_key(dd, id);
                        if (nkey >= key[Lq - 1]) continue;
                        /* InsertIntoPool, index.h:152-180, on (dist,id):
                         * position = number of smaller keys; an equal key is
                         * the same id -> duplicate (returns L+1). */
                        uint32_t pos = (uint32_t)(std::lower_bound(key.begin(), key.begin() + Lq, nkey) - key.begin());
                        uint32_t r;
                        if (key[pos] == nkey) {
                            r = Lq + 1;
                        } else {
                            std::memmove(&key[pos + 1], &key[pos], (Lq - 1 - pos) * sizeof(uint64_t));
                            std::memmove(&flag[pos + 1], &flag[pos], (Lq - 1 - pos));
                            key[pos] = nkey;
                            flag[pos] = 1;
                            r = pos;
                        }
                        if (r < nk) nk = r;
                        if (r <= kidx) st = hp;
                    }
                }
                kk = (nk <= kk) ? nk : kk + 1;
                if (expanded && watch) {
                    float f[8];
                    const float kth = key_dist(key[kidx]);
                    wvo_term_features((float)hp, kth, key_dist(key[0]), entry_d, hist[3], (float)(hp - st),
                                      (float)Lq, (float)kq, f);
                    hist[3] = hist[2];
                    hist[2] = hist[1];
                    hist[1] = hist[0];
                    hist[0] = kth;
                    if (trace && hp <= cap)
                        for (int i = 0; i < 8; ++i) trace[((size_t)q * cap + (hp - 1)) * 8 + i] = f[i];
                    /* the rule is evaluated after every 4th expansion */
                    if (model && hp % 4 == 0 && kk < Lq && hp >= model->min_hops &&
                        logit_bf16(model, f) > model->logit_threshold)
                        break;
                }
            }
            if (trace_len) trace_len[q] = hp < cap ? hp : cap;
            if (settle) settle[q] = st;
            for (uint32_t i = 0; i < k_stride; ++i) {
                out_ids[(size_t)q * k_stride + i] = i < kq ? key_id(key[i]) : 0xFFFFFFFFu;
                if (out_dist) out_dist[(size_t)q * k_stride + i] = i < kq ? key_dist(key[i]) : -1.0f;
            }
            if (evals) evals[q] = ev;
            if (hops) hops[q] = hp;
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Exact kNN (FP64)                                                        */
/* ---------------------------------------------------------------------- */
int wvo_exact_knn(const float *base, uint32_t N, const float *queries, uint32_t Q, uint32_t d,
                  uint32_t k, int64_t self_offset, uint32_t *out_ids, double *out_dist,
                  int nthreads) {
    if (!base || !queries || !out_ids || k == 0 || k > N) return -1;
    const int nth = nthreads_or_default(nthreads);
#pragma omp parallel num_threads(nth)
    {
        std::vector<std::pair<double, uint32_t>> h;
#pragma omp for schedule(dynamic, 4)
        for (int64_t qq = 0; qq < (int64_t)Q; ++qq) {
            const float *qv = queries + (size_t)qq * d;
            int64_t self = self_offset >= 0 ? qq + self_offset : -1;
            h.clear();
            for (uint32_t j = 0; j < N; ++j) {
                if ((int64_t)j == self) continue;
                std::pair<double, uint32_t> e(wvo_dist_f64(qv, base + (size_t)j * d, d), j);
                if (h.size() < k) {
                    h.push_back(e);
                    std::push_heap(h.begin(), h.end());
                } else if (e < h.front()) {
                    std::pop_heap(h.begin(), h.end());
                    h.back() = e;
                    std::push_heap(h.begin(), h.end());
                }
            }
            std::sort_heap(h.begin(), h.end());
            for (uint32_t i = 0; i < k; ++i) {
                out_ids[(size_t)qq * k + i] = i < h.size() ? h[i].second : 0xFFFFFFFFu;
                if (out_dist) out_dist[(size_t)qq * k + i] = i < h.size() ? h[i].first : -1.0;
            }
        }
    }
    return 0;
}

double wvo_recall(const uint32_t *res, uint32_t Q, uint32_t k, uint32_t res_stride,
                  const uint32_t *gt, uint32_t gt_dim) {
    /* src/builder.cpp:465-480: count result ids that are not among the first
     * k ground-truth ids; accuracy = 1 - misses / (Q * k). */
    if (Q == 0 || k == 0 || k > gt_dim) return 0.0;
    uint64_t miss = 0;
    for (uint32_t i = 0; i < Q; ++i)
        for (uint32_t j = 0; j < k; ++j) {
            uint32_t r = res[(size_t)i * res_stride + j];
            bool hit = false;
            for (uint32_t t = 0; t < k; ++t)
                if (gt[(size_t)i * gt_dim + t] == r) { hit = true; break; }
            miss += !hit;
        }
    return 1.0 - (double)miss / ((double)Q * k);
}

double wvo_graph_quality(const uint32_t *ids, const uint32_t *deg, uint32_t deg_stride,
                         const uint32_t *rows, uint32_t nrows, const uint32_t *gt, uint32_t gt_dim,
                         uint32_t control) {
    /* src/builder.cpp:760-776: for each node, the fraction of its first
     * `control` exact neighbours present anywhere in its adjacency list,
     * normalised by min(control, deg, |exact|); averaged over nodes.  gt row
     * r belongs to node rows[r] (rows NULL: r itself). */
    if (nrows == 0) return 0.0;
    double mean = 0.0;
    for (uint32_t r = 0; r < nrows; ++r) {
        uint32_t node = rows ? rows[r] : r;
        const uint32_t *g = ids + (size_t)node * deg_stride;
        uint32_t dg = deg[node];
        const uint32_t *v = gt + (size_t)r * gt_dim;
        uint32_t q = 0;
        for (uint32_t j = 0; j < control && j < gt_dim; ++j)
            for (uint32_t t = 0; t < dg; ++t)
                if (g[t] == v[j]) { ++q; break; }
        uint32_t den = std::min(std::min(control, dg), gt_dim);
        if (den) mean += (double)q / den;
    }
    return mean / nrows;
}

}  // extern "C"

/* ---------------------------------------------------------------------- */
/* NSG refine (src/component_refine.cpp:393-540, component_candidate.cpp:  */
/* 10-85, component_prune.cpp:20-66, component_conn.cpp:10-140)            */
/* ---------------------------------------------------------------------- */
namespace {
constexpr uint64_t TAG_NSGC = 0x4E534743ull;  /* "NSGC": candidate-search entries */
constexpr uint64_t TAG_NSGR = 0x4E534752ull;  /* "NSGR": connectivity fallback root */
constexpr uint32_t NSG_TCAP = 1024;           /* candidates of one reverse-insert re-prune */

using Adj = std::vector<std::vector<uint32_t>>;

/* ep's first min(L, degree) distinct in-range ids: the head of every
 * candidate search's entries. */
std::vector<uint32_t> nsg_head(const Adj &g, uint32_t ep, uint32_t L, uint32_t N) {
    std::vector<uint32_t> h;
    for (uint32_t x : g[ep]) {
        if (h.size() == L) break;
        if (x < N && std::find(h.begin(), h.end(), x) == h.end()) h.push_back(x);
    }
    return h;
}

/* Entries of node n's candidate search: the head, then the ids of
 * GenRandom(L) on n's "NSGC" stream that are not in the head, in order, up
 * to L (the reference fills with rand(), component_candidate.cpp:17-31). */
void nsg_entries(const std::vector<uint32_t> &head, const std::vector<uint32_t> &head_sorted, uint32_t L,
                 uint32_t N, uint64_t seed, uint32_t n, std::vector<uint32_t> &out) {
    out.assign(head.begin(), head.end());
    std::vector<uint32_t> gen(L);
    gen_random(stream_of(seed, TAG_NSGC, n), gen.data(), L, N);
    for (uint32_t x : gen) {
        if (out.size() == L) break;
        if (!std::binary_search(head_sorted.begin(), head_sorted.end(), x)) out.push_back(x);
    }
}

/* CandidateInner (component_candidate.cpp:10-85): greedy search for point
 * qv from `ent` with pool L, as RouteGreedy; returns the C smallest
 * distinct keys of everything evaluated, entries included. */
void nsg_collect(const Adj &g, const float *base, uint32_t d, uint32_t N, const float *qv,
                 const std::vector<uint32_t> &ent, uint32_t L, uint32_t C, std::vector<uint32_t> &stamp,
                 uint32_t &epoch, std::vector<uint64_t> &out) {
    if (++epoch == 0) {
        std::fill(stamp.begin(), stamp.end(), 0);
        epoch = 1;
    }
    std::vector<uint64_t> key(L + 1, KEY_EMPTY), all;
    std::vector<uint8_t> flag(L + 1, 0);
    for (uint32_t i = 0; i < L; ++i) {
        key[i] = make_key(dist_canon(qv, base + (size_t)ent[i] * d, d), ent[i]);
        flag[i] = 1;
        all.push_back(key[i]);
    }
    std::sort(key.begin(), key.begin() + L);
    uint32_t kk = 0;
    while (kk < L) {
        uint32_t nk = L;
        if (flag[kk]) {
            flag[kk] = 0;
            for (uint32_t id : g[key_id(key[kk])]) {
                if (id >= N || stamp[id] == epoch) continue;
                stamp[id] = epoch;
                const uint64_t nkey = make_key(dist_canon(qv, base + (size_t)id * d, d), id);
                all.push_back(nkey);
                if (nkey >= key[L - 1]) continue;
                const uint32_t pos = (uint32_t)(std::lower_bound(key.begin(), key.begin() + L, nkey) - key.begin());
                if (key[pos] == nkey) continue;
                std::memmove(&key[pos + 1], &key[pos], (L - 1 - pos) * sizeof(uint64_t));
                std::memmove(&flag[pos + 1], &flag[pos], L - 1 - pos);
                key[pos] = nkey;
                flag[pos] = 1;
                if (pos < nk) nk = pos;
            }
        }
        kk = (nk <= kk) ? nk : kk + 1;
    }
    std::sort(all.begin(), all.end());
    all.erase(std::unique(all.begin(), all.end()), all.end());
    if (all.size() > C) all.resize(C);
    out.swap(all);
}

/* How a refiner turns candidates into a list: NSG (component_prune.cpp:
 * 20-66) adds the node's kNN neighbours and occludes with alpha = 1;
 * Vamana (component_prune.cpp:245-297) uses the candidates alone, occludes
 * with its alpha and fills up to R with the nearest occluded ones. */
struct Rule {
    bool add_knn, fill, conn;
    float alpha;
};

/* The occlusion rule over sorted distinct keys c[0..m) of node `self`: self
 * skipped; each next key is kept unless a kept r has alpha * d(r, key) <
 * d(node, key); at most R.  Written as: every kept key removes the later
 * keys it occludes.  With `fill`, a list shorter than R is topped up with
 * the nearest removed keys, in order. */
void nsg_occlude(const float *base, uint32_t d, const uint64_t *c, uint32_t m, uint32_t self, uint32_t R,
                 std::vector<uint64_t> &res, float alpha = 1.0f, bool fill = false) {
    res.clear();
    std::vector<uint8_t> alive(m, 1);
    for (uint32_t i = 0; i < m && res.size() < R; ++i) {
        if (!alive[i] || key_id(c[i]) == self) continue;
        res.push_back(c[i]);
        alive[i] = 2;
        const float *ri = base + (size_t)key_id(c[i]) * d;
        for (uint32_t j = i + 1; j < m; ++j)
            if (alive[j] == 1 && alpha * dist_canon(ri, base + (size_t)key_id(c[j]) * d, d) < key_dist(c[j]))
                alive[j] = 0;
    }
    if (fill)
        for (uint32_t i = 0; i < m && res.size() < R; ++i)
            if (alive[i] == 0 && key_id(c[i]) != self) res.push_back(c[i]);
}

/* Link (component_refine.cpp:438-472) without the connectivity pass:
 * candidates, prune, then InterInsert (:474-540) as one batch: node t's
 * list grows by every s with t in s's pruned list (ascending s, skipping
 * ones already there); a list above R is sorted, cut to its NSG_TCAP
 * nearest and re-pruned. */
std::vector<std::vector<uint64_t>> nsg_link(const Adj &g, const float *base, uint32_t N, uint32_t d, uint32_t ep,
                                            uint32_t L, uint32_t R, uint32_t C, uint64_t seed, int nth,
                                            const Rule &rule) {
    const std::vector<uint32_t> head = nsg_head(g, ep, L, N);
    std::vector<uint32_t> hs(head);
    std::sort(hs.begin(), hs.end());
    std::vector<std::vector<uint64_t>> pr(N);
#pragma omp parallel num_threads(nth)
    {
        std::vector<uint32_t> stamp(N, 0), ent;
        uint32_t epoch = 0;
        std::vector<uint64_t> F;
#pragma omp for schedule(dynamic, 16)
        for (int64_t nn = 0; nn < (int64_t)N; ++nn) {
            const uint32_t n = (uint32_t)nn;
            const float *qv = base + (size_t)n * d;
            nsg_entries(head, hs, L, N, seed, n, ent);
            nsg_collect(g, base, d, N, qv, ent, L, C, stamp, epoch, F);
            if (rule.add_knn)  /* PruneNSG adds the node's own neighbours (:26-34) */
                for (uint32_t id : g[n])
                    if (id < N) F.push_back(make_key(dist_canon(qv, base + (size_t)id * d, d), id));
            std::sort(F.begin(), F.end());
            F.erase(std::unique(F.begin(), F.end()), F.end());
            if (F.size() > C) F.resize(C);
            nsg_occlude(base, d, F.data(), (uint32_t)F.size(), n, R, pr[n], rule.alpha, rule.fill);
        }
    }
    std::vector<std::vector<uint64_t>> inc(N);
    for (uint32_t n = 0; n < N; ++n)
        for (uint64_t k : pr[n]) inc[key_id(k)].push_back(make_key(key_dist(k), n));
    std::vector<std::vector<uint64_t>> out(N);
#pragma omp parallel for num_threads(nth) schedule(dynamic, 64)
    for (int64_t tt = 0; tt < (int64_t)N; ++tt) {
        const uint32_t t = (uint32_t)tt;
        std::vector<uint64_t> temp(pr[t]);
        for (uint64_t k : inc[t]) {
            bool dup = false;
            for (uint64_t o : pr[t])
                if (key_id(o) == key_id(k)) { dup = true; break; }
            if (!dup) temp.push_back(k);
        }
        if (temp.size() <= R) {
            out[t].swap(temp);
        } else {
            std::sort(temp.begin(), temp.end());
            if (temp.size() > NSG_TCAP) temp.resize(NSG_TCAP);
            nsg_occlude(base, d, temp.data(), (uint32_t)temp.size(), t, R, out[t], rule.alpha, rule.fill);
        }
    }
    return out;
}

void nsg_mark(const std::vector<std::vector<uint64_t>> &G, uint32_t root, std::vector<uint8_t> &flag) {
    std::vector<uint32_t> st{root};
    flag[root] = 1;
    while (!st.empty()) {
        const uint32_t u = st.back();
        st.pop_back();
        for (uint64_t k : G[u]) {
            const uint32_t v = key_id(k);
            if (!flag[v]) {
                flag[v] = 1;
                st.push_back(v);
            }
        }
    }
}

/* tree_grow (component_conn.cpp:14-98): nodes unreachable from ep are linked
 * in ascending id order.  Each gets one candidate search on the graph as it
 * stands before any link (the reference searches the growing graph); its
 * root is the first reached node among the C nearest it evaluated, else the
 * first reached draw of its "NSGR" stream; the edge root -> node is
 * appended and the node's reach is marked. */
uint32_t nsg_conn(std::vector<std::vector<uint64_t>> &G, const float *base, uint32_t N, uint32_t d, uint32_t ep,
                  uint32_t L, uint32_t C, uint64_t seed, int nth) {
    std::vector<uint8_t> flag(N, 0);
    nsg_mark(G, ep, flag);
    std::vector<uint32_t> un;
    for (uint32_t i = 0; i < N; ++i)
        if (!flag[i]) un.push_back(i);
    if (un.empty()) return 0;
    Adj g(N);
    for (uint32_t i = 0; i < N; ++i)
        for (uint64_t k : G[i]) g[i].push_back(key_id(k));
    const std::vector<uint32_t> head = nsg_head(g, ep, L, N);
    std::vector<uint32_t> hs(head);
    std::sort(hs.begin(), hs.end());
    std::vector<std::vector<uint64_t>> F(un.size());
#pragma omp parallel num_threads(nth)
    {
        std::vector<uint32_t> stamp(N, 0), ent;
        uint32_t epoch = 0;
#pragma omp for schedule(dynamic, 4)
        for (int64_t i = 0; i < (int64_t)un.size(); ++i) {
            nsg_entries(head, hs, L, N, seed, un[i], ent);
            nsg_collect(g, base, d, N, base + (size_t)un[i] * d, ent, L, C, stamp, epoch, F[i]);
        }
    }
    uint32_t edges = 0;
    for (size_t i = 0; i < un.size(); ++i) {
        const uint32_t u = un[i];
        if (flag[u]) continue;
        uint32_t root = N;
        float dd = 0.f;
        for (uint64_t k : F[i])
            if (flag[key_id(k)]) {
                root = key_id(k);
                dd = key_dist(k);
                break;
            }
        if (root == N) {
            const uint64_t s = stream_of(seed, TAG_NSGR, u);
            for (uint64_t t = 0; t < (1ull << 26) && root == N; ++t) {
                const uint32_t x = draw(s, t) % N;
                if (flag[x]) root = x;
            }
            if (root == N) root = ep;
            dd = dist_canon(base + (size_t)root * d, base + (size_t)u * d, d);
        }
        G[root].push_back(make_key(dd, u));
        ++edges;
        nsg_mark(G, u, flag);
    }
    return edges;
}

Adj adj_from(const uint32_t *ids, const uint32_t *deg, uint32_t stride, uint32_t N) {
    Adj g(N);
    for (uint32_t i = 0; i < N; ++i) {
        const uint32_t c = std::min(deg[i], stride);
        g[i].assign(ids + (size_t)i * stride, ids + (size_t)i * stride + c);
    }
    return g;
}

int write_graph(const std::vector<std::vector<uint64_t>> &G, uint32_t *out_ids, float *out_dist, uint32_t *out_deg,
                uint32_t wcap, uint32_t *width) {
    uint32_t w = 1;
    for (const auto &r : G) w = std::max(w, (uint32_t)r.size());
    if (width) *width = w;
    if (w > wcap) return -2;
    for (size_t i = 0; i < G.size(); ++i) {
        out_deg[i] = (uint32_t)G[i].size();
        for (uint32_t j = 0; j < wcap; ++j) {
            const bool ok = j < G[i].size();
            out_ids[i * wcap + j] = ok ? key_id(G[i][j]) : 0xFFFFFFFFu;
            if (out_dist) out_dist[i * wcap + j] = ok ? key_dist(G[i][j]) : -1.0f;
        }
    }
    return 0;
}
}  // namespace

extern "C" {

int wvo_nsg(const float *base, uint32_t N, uint32_t d, const uint32_t *knn, const uint32_t *kdeg, uint32_t kstride,
            uint32_t L, uint32_t R, uint32_t C, uint32_t ep, uint64_t seed, uint32_t *out_ids, float *out_dist,
            uint32_t *out_deg, uint32_t wcap, uint32_t *width, uint32_t *conn_edges, int nthreads) {
    if (!base || !knn || !kdeg || !out_ids || !out_deg || N < 2 || d == 0) return -1;
    if (L == 0 || L >= N || R == 0 || C == 0 || ep >= N || kstride == 0) return -1;
    const int nth = nthreads_or_default(nthreads);
    const Adj g = adj_from(knn, kdeg, kstride, N);
    auto G = nsg_link(g, base, N, d, ep, L, R, C, seed, nth, Rule{true, false, true, 1.0f});
    const uint32_t e = nsg_conn(G, base, N, d, ep, L, C, seed, nth);
    if (conn_edges) *conn_edges = e;
    return write_graph(G, out_ids, out_dist, out_deg, wcap, width);
}

int wvo_vamana(const float *base, uint32_t N, uint32_t d, const uint32_t *ids, const uint32_t *deg, uint32_t stride,
               uint32_t L, uint32_t R, uint32_t C, float alpha, uint32_t ep, uint64_t seed, uint32_t *out_ids,
               float *out_dist, uint32_t *out_deg, uint32_t wcap, uint32_t *width, int nthreads) {
    if (!base || !ids || !deg || !out_ids || !out_deg || N < 2 || d == 0) return -1;
    if (L == 0 || L >= N || R == 0 || C == 0 || ep >= N || stride == 0 || !(alpha >= 1.0f)) return -1;
    const Adj g = adj_from(ids, deg, stride, N);
    auto G = nsg_link(g, base, N, d, ep, L, R, C, seed, nthreads_or_default(nthreads),
                      Rule{false, true, false, alpha});
    return write_graph(G, out_ids, out_dist, out_deg, wcap, width);
}

int wvo_dpg(const float *base, uint32_t N, uint32_t d, const uint32_t *ids, const uint32_t *deg, uint32_t stride,
            uint32_t range, uint32_t *out_ids, float *out_dist, uint32_t *out_deg, uint32_t wcap, uint32_t *width,
            int nthreads) {
    /* ComponentRefineDPG (src/component_refine.cpp:728-790): per node, its
     * list by (distance, id), the first len = min(size, 2 range); hit[j] =
     * the earlier entries i with d(p_i, p_j) < d(node, p_j); cut = the
     * (range+1)-th smallest hit; keep, in order, the entries with hit <= cut,
     * at most range (all of them when len <= range, where the reference reads
     * past its array).  Then ComponentConnReverse (component_conn.cpp:
     * 228-268): each list takes every reverse edge, sorted, repeated ids
     * dropped. */
    if (!base || !ids || !deg || !out_ids || !out_deg || N < 2 || d == 0 || range == 0 || stride == 0) return -1;
    const int nth = nthreads_or_default(nthreads);
    std::vector<std::vector<uint64_t>> pr(N);
#pragma omp parallel for num_threads(nth) schedule(dynamic, 64)
    for (int64_t nn = 0; nn < (int64_t)N; ++nn) {
        const uint32_t n = (uint32_t)nn;
        const float *x = base + (size_t)n * d;
        std::vector<uint64_t> p;
        for (uint32_t j = 0; j < std::min(deg[n], stride); ++j) {
            const uint32_t v = ids[(size_t)n * stride + j];
            if (v < N && v != n) p.push_back(make_key(dist_canon(x, base + (size_t)v * d, d), v));
        }
        std::sort(p.begin(), p.end());
        p.erase(std::unique(p.begin(), p.end()), p.end());
        const uint32_t len = std::min<uint32_t>((uint32_t)p.size(), 2 * range);
        if (len <= range) {
            pr[n].assign(p.begin(), p.begin() + len);
            continue;
        }
        std::vector<uint32_t> hit(len, 0);
        for (uint32_t i = 0; i + 1 < len; ++i)
            for (uint32_t j = i + 1; j < len; ++j)
                if (dist_canon(base + (size_t)key_id(p[i]) * d, base + (size_t)key_id(p[j]) * d, d) <
                    key_dist(p[j]))
                    ++hit[j];
        std::vector<uint32_t> sh(hit);
        std::sort(sh.begin(), sh.end());
        const uint32_t cut = sh[range];
        for (uint32_t j = 0; j < len && pr[n].size() < range; ++j)
            if (hit[j] <= cut) pr[n].push_back(p[j]);
    }
    std::vector<std::vector<uint64_t>> G(pr);
    for (uint32_t n = 0; n < N; ++n)
        for (uint64_t k : pr[n]) G[key_id(k)].push_back(make_key(key_dist(k), n));
#pragma omp parallel for num_threads(nth) schedule(dynamic, 64)
    for (int64_t t = 0; t < (int64_t)N; ++t) {
        std::sort(G[t].begin(), G[t].end());
        G[t].erase(std::unique(G[t].begin(), G[t].end()), G[t].end());
    }
    return write_graph(G, out_ids, out_dist, out_deg, wcap, width);
}

namespace {
constexpr uint64_t TAG_SSGE = 0x53534745ull;  /* "SSGE": SSG's n_try roots */

/* PruneSSG's angle rule (component_prune.cpp:71-122) over sorted distinct
 * keys c[0..m) of node `self` (key distance = squared distance to the node):
 * each next key p is kept unless a kept r has cos(p, r) > thr, with
 * cos = (d(p) + d(r) - d(r, p)) / 2 / sqrt(d(p) d(r)) in FP32 (law of
 * cosines on squared distances); at most R. */
void ssg_occlude(const float *base, uint32_t d, const uint64_t *c, uint32_t m, uint32_t self, uint32_t R, float thr,
                 std::vector<uint64_t> &res) {
    res.clear();
    std::vector<uint8_t> alive(m, 1);
    for (uint32_t i = 0; i < m && res.size() < R; ++i) {
        if (!alive[i] || key_id(c[i]) == self) continue;
        res.push_back(c[i]);
        const float *ri = base + (size_t)key_id(c[i]) * d;
        const float dr = key_dist(c[i]);
        for (uint32_t j = i + 1; j < m; ++j) {
            if (!alive[j]) continue;
            const float dp = key_dist(c[j]);
            const float djk = dist_canon(ri, base + (size_t)key_id(c[j]) * d, d);
            const float cs = ((dp + dr) - djk) / 2.0f / std::sqrt(dp * dr);
            if (cs > thr) alive[j] = 0;
        }
    }
}
}  // namespace

float wvo_ssg_threshold(float angle) {
    /* component_refine.cpp:638-639, component_prune.cpp:89-90: cos(A / 180 * pi), A in FP32, pi and
     * the cosine in FP64, rounded to FP32 */
    const double kPi = std::acos(-1.0);
    return (float)std::cos(angle / 180 * kPi);
}

int wvo_ssg(const float *base, uint32_t N, uint32_t d, const uint32_t *knn, const uint32_t *kdeg, uint32_t kstride,
            uint32_t L, uint32_t R, float angle, uint32_t n_try, uint64_t seed, uint32_t *out_ids, float *out_dist,
            uint32_t *out_deg, uint32_t wcap, uint32_t *width, uint32_t *eps, uint32_t *conn_edges, int nthreads) {
    /* ComponentRefineSSG (src/component_refine.cpp:550-727). */
    if (!base || !knn || !kdeg || !out_ids || !out_deg || N < 2 || d == 0 || L == 0 || R == 0) return -1;
    if (n_try == 0 || n_try >= N || kstride == 0) return -1;
    const int nth = nthreads_or_default(nthreads);
    const float thr = wvo_ssg_threshold(angle);
    const Adj g = adj_from(knn, kdeg, kstride, N);
    std::vector<std::vector<uint64_t>> pr(N);
#pragma omp parallel for num_threads(nth) schedule(dynamic, 64)
    for (int64_t nn = 0; nn < (int64_t)N; ++nn) {
        const uint32_t n = (uint32_t)nn;
        const float *x = base + (size_t)n * d;
        /* Propagation2 (component_candidate.cpp:89-108): the neighbours'
         * neighbours in list order, each once and never the node, until L */
        std::vector<uint32_t> seen{n};
        std::vector<uint64_t> c;
        for (uint32_t nid : g[n]) {
            if (c.size() >= L) break;
            if (nid >= N) continue;
            for (uint32_t v : g[nid]) {
                if (v >= N || std::find(seen.begin(), seen.end(), v) != seen.end()) continue;
                seen.push_back(v);
                c.push_back(make_key(dist_canon(x, base + (size_t)v * d, d), v));
                if (c.size() >= L) break;
            }
        }
        for (uint32_t v : g[n])  /* PruneSSG adds the node's own list (:74-81) */
            if (v < N && v != n) c.push_back(make_key(dist_canon(x, base + (size_t)v * d, d), v));
        std::sort(c.begin(), c.end());
        c.erase(std::unique(c.begin(), c.end()), c.end());
        ssg_occlude(base, d, c.data(), (uint32_t)c.size(), n, R, thr, pr[n]);
    }
    /* InterInsert with the angle rule (:647-727), as one batch */
    std::vector<std::vector<uint64_t>> inc(N);
    for (uint32_t n = 0; n < N; ++n)
        for (uint64_t k : pr[n]) inc[key_id(k)].push_back(make_key(key_dist(k), n));
    std::vector<std::vector<uint64_t>> G(N);
#pragma omp parallel for num_threads(nth) schedule(dynamic, 64)
    for (int64_t tt = 0; tt < (int64_t)N; ++tt) {
        const uint32_t t = (uint32_t)tt;
        std::vector<uint64_t> temp(pr[t]);
        for (uint64_t k : inc[t]) {
            bool dup = false;
            for (uint64_t o : pr[t])
                if (key_id(o) == key_id(k)) { dup = true; break; }
            if (!dup) temp.push_back(k);
        }
        if (temp.size() <= R) {
            G[t].swap(temp);
        } else {
            std::sort(temp.begin(), temp.end());
            if (temp.size() > NSG_TCAP) temp.resize(NSG_TCAP);
            ssg_occlude(base, d, temp.data(), (uint32_t)temp.size(), t, R, thr, G[t]);
        }
    }
    /* ComponentConnSSGDFS (component_conn.cpp:169-226): roots eps = GenRandom
     * (n_try) on the "SSGE" stream, taken in order; from each, a BFS; when
     * it stops short, the smallest unreached node u gets an edge from the
     * smallest reached node with fewer than R neighbours (if any), u counts
     * as reached, and the BFS goes on from u. */
    std::vector<uint32_t> roots(n_try);
    gen_random(stream_of(seed, TAG_SSGE, 0), roots.data(), n_try, N);
    uint32_t edges = 0;
    for (uint32_t r = 0; r < n_try; ++r) {
        std::vector<uint8_t> flag(N, 0);
        std::vector<uint32_t> q{roots[r]};
        flag[roots[r]] = 1;
        size_t head = 0;
        uint32_t low = 0;  /* every id below `low` is reached */
        for (;;) {
            while (head < q.size()) {
                const uint32_t u = q[head++];
                for (uint64_t k : G[u]) {
                    const uint32_t v = key_id(k);
                    if (!flag[v]) {
                        flag[v] = 1;
                        q.push_back(v);
                    }
                }
            }
            while (low < N && flag[low]) ++low;
            if (low == N) break;
            const uint32_t u = low;
            for (uint32_t j = 0; j < N; ++j)
                if (flag[j] && G[j].size() < R) {
                    G[j].push_back(make_key(dist_canon(base + (size_t)j * d, base + (size_t)u * d, d), u));
                    ++edges;
                    break;
                }
            flag[u] = 1;
            q.push_back(u);
        }
    }
    if (eps) std::memcpy(eps, roots.data(), n_try * 4);
    if (conn_edges) *conn_edges = edges;
    return write_graph(G, out_ids, out_dist, out_deg, wcap, width);
}

int wvo_nsg_conn(const float *base, uint32_t N, uint32_t d, const uint32_t *ids, const uint32_t *deg,
                 uint32_t stride, uint32_t L, uint32_t C, uint32_t ep, uint64_t seed, uint32_t *out_ids,
                 float *out_dist, uint32_t *out_deg, uint32_t wcap, uint32_t *width, uint32_t *conn_edges,
                 int nthreads) {
    if (!base || !ids || !deg || !out_ids || !out_deg || N < 2 || L == 0 || L >= N || C == 0 || ep >= N) return -1;
    std::vector<std::vector<uint64_t>> G(N);
    for (uint32_t i = 0; i < N; ++i)
        for (uint32_t j = 0; j < std::min(deg[i], stride); ++j) {
            const uint32_t v = ids[(size_t)i * stride + j];
            if (v < N) G[i].push_back(make_key(dist_canon(base + (size_t)i * d, base + (size_t)v * d, d), v));
        }
    const uint32_t e = nsg_conn(G, base, N, d, ep, L, C, seed, nthreads_or_default(nthreads));
    if (conn_edges) *conn_edges = e;
    return write_graph(G, out_ids, out_dist, out_deg, wcap, width);
}

}  // extern "C"
