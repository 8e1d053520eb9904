/*
 * weavess_hip.h -- C ABI of libweavess_hip.so, the MI355X (gfx950) kNN-graph
 * hot path of WEAVESS: NN-Descent graph construction and greedy best-first
 * search.
 *
 * Plain C, POD arguments, no exceptions and no torch types across the
 * boundary.  Every call returns WV_OK (0) or a negative WV_E* code;
 * wv_last_error() gives the text of the last failure on that context.
 * Calls on one context are serialised by an internal mutex.  All calls are
 * synchronous: when one returns, its outputs are written.  The caller owns
 * every host and device buffer it passes in.
 *
 * Reference interfaces replaced (paths relative to the WEAVESS tree):
 *   wv_index_build_nndescent / wv_nndescent_build
 *       ComponentInitRandom::InitInner      src/component_init.cpp:11-45
 *     + ComponentRefineNNDescent::RefineInner src/component_refine.cpp:12-77
 *     + ComponentPruneNaive::PruneInner      src/component_prune.cpp:8-18
 *       (virtual hooks: include/weavess/component.h:32-37 InitInner,
 *        :361-366 RefineInner)
 *   wv_index_search / wv_search_greedy
 *       ComponentSearchEntryRand::SearchEntryInner src/component_search_entry.cpp:9-30
 *     + ComponentSearchRouteGreedy::RouteInner     src/component_route.cpp:9-53
 *       (virtual hooks: include/weavess/component.h:752-757, :827-832),
 *       batched over the query loop of IndexBuilder::search
 *       (src/builder.cpp:440-458)
 *   wv_exact_knn
 *       ground-truth generation, algo/knn/gen_knn.cpp:117-150 and
 *       dataset/main.cpp:237-270 (FP64, self excluded, ties by id)
 *
 * Ordering convention: every list is ascending in (distance, id).  The
 * reference orders by distance only; ids break its ties here.
 */
#ifndef WEAVESS_HIP_H
#define WEAVESS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WV_OK 0
#define WV_EINVAL (-22)    /* bad argument: null pointer, L < k, S >= N, ... */
#define WV_ENOMEM (-12)    /* device allocation failed */
#define WV_EHIP (-5)       /* a HIP runtime call failed */
#define WV_ENOTSUP (-95)   /* shape outside what the kernels support */

typedef struct wv_ctx wv_ctx;
typedef struct wv_index wv_index;

/* Join precision of NN-Descent. */
#define WV_DTYPE_F32 0
#define WV_DTYPE_BF16 1
#define WV_DTYPE_FP8 2

typedef struct {
    uint32_t K;          /* "K": final degree (src/component_refine.cpp:81)          */
    uint32_t L;          /* "L": pool size (:80)                                     */
    uint32_t S;          /* "S": initial and per-iteration new samples (component_init.cpp:47-49) */
    uint32_t R;          /* "R": reverse-sample cap (:83)                            */
    uint32_t iters;      /* "ITER": (join, update) rounds incl. the empty first join (:84, :121-124) */
    uint32_t dtype;      /* WV_DTYPE_*: distance precision inside the local join    */
    uint32_t rerank_mult;/* low precision: the best rerank_mult*K of each pool are re-ranked with
                            exact FP32 distances at the end (0 -> 2)                */
    uint32_t cand_slots; /* per-node candidate table, power of two in [64, 1024]
                            (0 -> automatic)                                        */
    uint64_t seed;       /* random initialisation and reverse-sample choice          */
    uint32_t order;      /* node processing order: 0 automatic (grouped from 32,768 nodes),
                            1 by id, 2 grouped by nearest pivot row.  Speed only: the
                            graph is the same in every order                         */
    uint32_t reserved;   /* 0                                                        */
} wv_nnd_params;

typedef struct {
    double total_ms;     /* InitRandom + NN-Descent + finalize, device time          */
    double init_ms, sample_ms, join_ms, merge_ms, finalize_ms;
    uint64_t pairs;      /* distance evaluations in all joins                        */
    uint64_t proposals;  /* join results that passed the target's threshold and were
                            not already in its pool (candidate-slot updates)         */
    uint64_t rows_gathered; /* sample rows loaded by the joins                       */
    uint32_t iters_run;
    uint32_t cand_slots;
} wv_build_stats;

typedef struct {
    double kernel_ms;    /* device time of the search launch(es)                     */
    uint64_t dist_evals; /* sum over queries (entries + first-visited neighbours)    */
    uint64_t hops;       /* sum over queries (expanded pool entries)                 */
} wv_search_stats;

/* Graph in "padded adjacency" form: ids[N*deg_max], unused slots any value,
 * deg[N] valid counts.  The reference's on-disk graph (u32 degree, then that
 * many u32 ids per node; src/builder.cpp:1068-1076) loads into this form. */
typedef struct {
    const uint32_t *ids;
    const uint32_t *deg;
    uint32_t N;
    uint32_t deg_max;
} wv_graph;

/* Search tuning (the autotuner's knobs; ids and distances never depend on
 * them, only speed and the evaluation count: a visited table too small to
 * hold every visited node forgets some, which are then evaluated again). */
typedef struct {
    uint32_t waves_per_block;  /* 1, 2 or 4 queries per workgroup (0 -> auto)        */
    uint32_t row_groups;       /* gather tile: 4*row_groups rows in flight (0 -> auto) */
    uint32_t visited_slots;    /* LDS visited table size, power of two >= 64
                                  (0 -> up to L = 256: 8L rounded up to a power
                                  of two, clamped to [1024, 2048]; up to 512:
                                  1024; above: 256)                                  */
    uint32_t merge_buffer;     /* keys buffered before a pool merge: 0 auto (256 from
                                  L = 2560 up), 1 none, or 64/128/192/256           */
    uint32_t schedule;         /* 0 or 1: one wavefront per query (the hardware
                                  dispatcher hands out workgroups); 2: persistent
                                  wavefronts, about as many as the GPU holds,
                                  taking queries from a device work queue, the
                                  largest per-query L first                          */
    uint32_t waves_per_query;  /* 0 or 1: one wavefront per query; 2 or 4: that many
                                  wavefronts of one workgroup share each query
                                  (the neighbours' visited tests and distances
                                  split between them); plain search only, the
                                  learned stop and traces run one wavefront      */
} wv_search_config;

/* Learned early termination: a 2-layer MLP (8 -> 32 ReLU -> 1) over
 * per-hop search-state features, evaluated in BF16 (FP32 accumulation)
 * after every 4th expansion.  Search stops when the logit exceeds
 * logit_threshold and at least min_hops expansions were made.  (Traces
 * record the features after every expansion.)  Features
 * g(state), standardised as (g - mean) * inv_std:
 *   0 log2(1 + hops)          1 kth / entry     2 best / entry
 *   3 log2(entry)             4 (kth 4 hops ago - kth) / entry
 *   5 log2(L)                 6 log2(k)         7 log2(1 + stale)
 * where entry = best entry distance, kth/best = current k-th/best pool
 * distance (squared L2, the library's distances; entry is clamped to
 * >= 1e-30), "kth 4 hops ago" is the entries' k-th distance for the first
 * 4 expansions, and stale = expansions since a candidate last entered the
 * top-k. */
typedef struct {
    float w1[32][8];
    float b1[32];
    float w2[32];
    float b2;
    float mean[8];
    float inv_std[8];
    float logit_threshold;
    uint32_t min_hops;
} wv_term_model;

/* Optional extras of wv_index_search_ext[_dev]. */
typedef struct {
    const wv_term_model *model;  /* learned termination (NULL: plain fixed-L)      */
    float *trace;                /* Q x trace_cap x 8 features g after each hop,
                                    or NULL (same memory space as the outputs;
                                    16-byte aligned)                                 */
    uint32_t *trace_len;         /* Q: rows written (min(hops, trace_cap))          */
    uint32_t *settle_hop;        /* Q: expansions after which the top-k never
                                    changed again (0: the entries' top-k is final)  */
    uint32_t trace_cap;
} wv_search_ext;

/* MLP training (FP32, Adam, mean binary cross-entropy on logits). */
typedef struct {
    uint32_t epochs;
    uint32_t batch;              /* multiple of 64                                   */
    float lr;
    float beta1, beta2, eps;
    float pos_weight;            /* weight of label-1 samples in the loss            */
    uint64_t seed;               /* minibatch order                                   */
} wv_train_params;

typedef struct {
    double first_loss, last_loss;  /* mean loss of the first and last epoch          */
    double kernel_ms;
    uint32_t steps;
} wv_train_stats;

/* ---------------------------------------------------------------- context */
int wv_ctx_create(int device, wv_ctx **out);
int wv_ctx_destroy(wv_ctx *ctx);
const char *wv_last_error(const wv_ctx *ctx);
int wv_version(void);

/* ---------------------------------------------------------------- index */
/* A device-resident base set (and, once built or set, its graph). */
int wv_index_create(wv_ctx *ctx, const float *base, uint32_t N, uint32_t d, wv_index **out);
int wv_index_destroy(wv_index *idx);
/* Replace the graph (host arrays, padded adjacency form). */
int wv_index_set_graph(wv_index *idx, const uint32_t *ids, const uint32_t *deg, uint32_t deg_max);
/* Copy the graph to the host: ids/dist are N*deg_max (dist may be NULL). */
int wv_index_get_graph(wv_index *idx, uint32_t *ids, float *dist, uint32_t *deg, uint32_t *deg_max);
int wv_index_set_search_config(wv_index *idx, const wv_search_config *cfg);

/* NN-Descent on the resident base; the result becomes the index's graph
 * (deg_max = K).  init_ids: optional N*S initial neighbour ids (host), as
 * ComponentRefineNNDescent::init reads them from FinalGraph
 * (src/component_refine.cpp:96-103); NULL runs InitRandom on the device. */
int wv_index_build_nndescent(wv_index *idx, const wv_nnd_params *p, const uint32_t *init_ids,
                             const uint32_t *init_cnt, wv_build_stats *stats);

/* wv_index_build_nndescent with the per-round graph-recall monitor of the
 * reference's debug path (ComponentRefineNNDescent::NNDescent with
 * index->debug, src/component_refine.cpp:106-118; eval_recall :258-275):
 * after round r (its join and update), per_round[r] = mean over the n_ctrl
 * control nodes of |pool ∩ truth row| / truth_k.  ctrl: n_ctrl node ids
 * (< N); truth: n_ctrl x truth_k exact neighbour ids, the caller's choice
 * (the reference takes 100 GenRandom nodes and their exact top 100
 * including the node itself, generate_control_set :241-256; wv_exact_knn
 * gives them).  per_round has p->iters entries.  1 <= n_ctrl <= 2^20,
 * 1 <= truth_k <= 1024.  The monitor's kernels run inside the build's
 * timed span. */
int wv_index_build_nndescent_monitor(wv_index *idx, const wv_nnd_params *p, const uint32_t *init_ids,
                                     const uint32_t *init_cnt, const uint32_t *ctrl, uint32_t n_ctrl,
                                     const uint32_t *truth, uint32_t truth_k, double *per_round,
                                     wv_build_stats *stats);

/* EFANNA's NN-Descent (ComponentRefineEFANNA, src/component_refine.cpp:
 * 959-1183) on the resident base; the result becomes the index's graph.
 * init_ids: optional N*init_stride ids of an initial graph (host; rows of
 * init_cnt[i] valid ids, or all init_stride when init_cnt is NULL;
 * 1 <= init_stride <= L), as EFANNA's init reads FinalGraph (:1036-1058):
 * each row is sorted by distance with self, out-of-range and repeated ids
 * dropped.  NULL starts from InitRandom's S random neighbours.  Unlike
 * KGraph's refine, the first join runs over 2S GenRandom ids per node
 * (include/weavess/index.h:105-110), so it needs 2S < N. */
int wv_index_build_efanna(wv_index *idx, const wv_nnd_params *p, const uint32_t *init_ids,
                          const uint32_t *init_cnt, uint32_t init_stride, wv_build_stats *stats);

/* NSG refine parameters (ComponentRefineNSG::SetConfigs,
 * src/component_refine.cpp:430-436). */
typedef struct {
    uint32_t L;     /* "L_refine": pool of each node's candidate search, 1..min(N-1, 1024) */
    uint32_t R;     /* "R_refine": degree bound of the pruned lists, 1..256       */
    uint32_t C;     /* "C_refine": candidates the prune considers, 1..1024         */
    uint32_t ep;    /* navigating node (ComponentRefineEntryCentroid's ep_)        */
    uint64_t seed;  /* random entry fill and fallback roots                        */
} wv_nsg_params;

typedef struct {
    /* stage times; conn_ms covers the connectivity links and storing the
     * result rows in the index */
    double total_ms, search_ms, prune_ms, insert_ms, conn_ms;
    uint64_t evals;       /* distance evaluations of the candidate searches */
    uint32_t conn_edges;  /* edges added to reach every node from ep        */
    uint32_t width;       /* widest row of the result                       */
} wv_nsg_stats;

/* NSG refine of the index's graph, which must be its kNN graph
 * (ComponentRefineNSG::RefineInner, src/component_refine.cpp:393-428, after
 * the centroid entry): per node a greedy search for the node from ep's
 * neighbours plus random ids gathers candidates
 * (ComponentCandidateNSG, component_candidate.cpp:10-85), PruneNSG's
 * occlusion keeps up to R of the C nearest (component_prune.cpp:20-66),
 * reverse edges are inserted and re-pruned (InterInsert, :474-540, as one
 * batch), and nodes unreachable from ep are linked (ComponentConnNSGDFS,
 * component_conn.cpp:10-140).  The result replaces the index's graph; rows
 * hold at most R ids plus the links. */
int wv_index_build_nsg(wv_index *idx, const wv_nsg_params *p, wv_nsg_stats *stats);

/* One pass of ComponentRefineVAMANA (src/component_refine.cpp:794-958): the
 * reference runs two, alpha 1 then 2 (:871), from INIT_RAND's random graph. */
typedef struct {
    uint32_t L;      /* "L_refine": pool of each node's candidate search        */
    uint32_t R;      /* "R_refine": list length                                  */
    uint32_t C;      /* candidates the prune considers, 1..1024                   */
    uint32_t ep;     /* navigating node                                           */
    float alpha;     /* occlusion factor, >= 1 (on squared distances)             */
    uint32_t reserved;
    uint64_t seed;
} wv_vamana_params;

/* Vamana pass over the index's graph: NSG's candidate search from ep, then
 * Vamana's prune (PruneVAMANA, component_prune.cpp:245-297): candidates by
 * distance, p kept unless a kept r has alpha * d(r, p) < d(node, p), at most
 * R, then topped up to R with the nearest dropped ones; reverse edges are
 * inserted with the same prune (InterInsert, :874-958, as one batch).  No
 * connectivity pass.  The result replaces the index's graph. */
int wv_index_build_vamana(wv_index *idx, const wv_vamana_params *p, wv_nsg_stats *stats);

/* SSG refine parameters (ComponentRefineSSG::SetConfigs,
 * src/component_refine.cpp:584-591). */
typedef struct {
    uint32_t L;      /* "L_refine": two-hop candidates a node, 1..2048             */
    uint32_t R;      /* "R_refine": list length, 1..256                             */
    float angle;     /* "A": minimum angle between kept edges, degrees (60)         */
    uint32_t n_try;  /* "n_try": roots of the connectivity pass (the entry set eps_) */
    uint64_t seed;   /* the roots                                                   */
} wv_ssg_params;

/* SSG refine of the index's graph, which must be its kNN graph
 * (ComponentRefineSSG::RefineInner, src/component_refine.cpp:550-582):
 * candidates are a node's neighbours' neighbours, up to L, plus its own
 * list (ComponentCandidatePropagation2, component_candidate.cpp:89-108); a
 * candidate is kept unless it makes an angle below `angle` with a kept one
 * (ComponentPruneSSG, component_prune.cpp:71-122); reverse edges are
 * inserted with the same rule (:647-727, as one batch); from n_try random
 * roots (written to eps) every node is linked in
 * (ComponentConnSSGDFS, component_conn.cpp:169-226).  The result replaces
 * the index's graph. */
int wv_index_build_ssg(wv_index *idx, const wv_ssg_params *p, uint32_t *eps, wv_nsg_stats *stats);

/* DPG refine of the index's graph, which must be its kNN graph
 * (ComponentRefineDPG, src/component_refine.cpp:728-790): each node keeps up
 * to `range` (the reference's "K" / 2) of its first 2 * range neighbours,
 * those occluded by the fewest nearer ones (ComponentPruneDPG,
 * component_prune.cpp:124-186); then every edge is mirrored and each list
 * sorted (ComponentConnReverse, component_conn.cpp:228-268).  Rows are not
 * bounded; stats->width reports the widest.  The result replaces the
 * index's graph. */
int wv_index_build_dpg(wv_index *idx, uint32_t range, wv_nsg_stats *stats);

/* Greedy search of Q queries (host arrays).  entries: Q*L ids or NULL
 * (seeded GenRandom per query).  L_per_q/k_per_q: NULL or Q values each
 * (then `entries` must be NULL).  Requires k <= L, and L < N for random
 * entries.  Outputs Q*k_stride (slots >= k get 0xFFFFFFFF / -1). */
int wv_index_search(wv_index *idx, const float *queries, uint32_t Q, const uint32_t *entries,
                    uint64_t seed, uint32_t L, uint32_t k, const uint32_t *L_per_q,
                    const uint32_t *k_per_q, uint32_t k_stride, uint32_t *out_ids,
                    float *out_dist, uint32_t *evals, uint32_t *hops, wv_search_stats *stats);
/* Same with device pointers.  queries: Q rows of stride roundup(d,4) floats,
 * zero padded.  Everything else as above, on the device (d_L_per_q,
 * d_k_per_q and d_entries are copied back once to check them). */
int wv_index_search_dev(wv_index *idx, const float *d_queries, uint32_t Q,
                        const uint32_t *d_entries, uint64_t seed, uint32_t L, uint32_t k,
                        const uint32_t *d_L_per_q, const uint32_t *d_k_per_q, uint32_t k_stride,
                        uint32_t *d_out_ids, float *d_out_dist, uint32_t *d_evals,
                        uint32_t *d_hops, wv_search_stats *stats);

/* Greedy search that allocates nothing, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  It
 * is wv_index_search_dev with seeded random entries and one L and k for
 * every query (no explicit entries, no per-query budgets), and returns the
 * same ids, distances, counts and statistics.  The per-query evaluation and
 * hop counts behind `stats`, when d_evals/d_hops are NULL, and the totals
 * live in a file-scope device array of the library, about 2 MiB, which each
 * GPU the library runs on holds until the process ends; Q above
 * WV_NOALLOC_MAX_Q is rejected with WV_EINVAL.  Every call on one GPU
 * shares that GPU's array: the library serialises such calls in one
 * process, and kernels from two calls that overlapped would race on it. */
#define WV_NOALLOC_MAX_Q (1u << 18)
int wv_index_search_dev_noalloc(wv_index *idx, const float *d_queries, uint32_t Q, uint64_t seed,
                                uint32_t L, uint32_t k, uint32_t k_stride, uint32_t *d_out_ids,
                                float *d_out_dist, uint32_t *d_evals, uint32_t *d_hops,
                                wv_search_stats *stats);

/* Greedy search with learned termination and/or per-hop traces. */
int wv_index_search_ext(wv_index *idx, const float *queries, uint32_t Q, const uint32_t *entries,
                        uint64_t seed, uint32_t L, uint32_t k, const uint32_t *L_per_q,
                        const uint32_t *k_per_q, uint32_t k_stride, uint32_t *out_ids,
                        float *out_dist, uint32_t *evals, uint32_t *hops,
                        const wv_search_ext *ext, wv_search_stats *stats);
int wv_index_search_ext_dev(wv_index *idx, const float *d_queries, uint32_t Q,
                            const uint32_t *d_entries, uint64_t seed, uint32_t L, uint32_t k,
                            const uint32_t *d_L_per_q, const uint32_t *d_k_per_q,
                            uint32_t k_stride, uint32_t *d_out_ids, float *d_out_dist,
                            uint32_t *d_evals, uint32_t *d_hops, const wv_search_ext *ext,
                            wv_search_stats *stats);

/* Search autotuner: times the search of Q sample queries (device rows as
 * for wv_index_search_dev; random entries from `seed`, one L and k, and the
 * model of `ext` when given) under candidate tunings, by coordinate descent
 * from the default plan over visited_slots {256..4096}, row_groups {4,8,16},
 * merge_buffer {none,128,256}, waves_per_block {1,2,4}, schedule
 * {per query, persistent} and waves_per_query {1,2,4}, each candidate
 * timed as the best of 3 launches after a warm-up and taken only when 2 %
 * faster than the current choice.  The fastest tuning is
 * kept in the index (as wv_index_set_search_config would) and written to
 * *best with its kernel time.  Results never depend on the tuning. */
int wv_index_autotune(wv_index *idx, const float *d_queries, uint32_t Q, uint64_t seed, uint32_t L,
                      uint32_t k, const wv_search_ext *ext, wv_search_config *best,
                      double *best_ms);
/* The same autotuner timing the search of the Q sample queries under their
 * own per-query budgets (d_L_per_q, d_k_per_q: Q values each on the device,
 * checked as for wv_index_search_dev; outputs of stride k_stride), for
 * launches that serve uneven L and k: the schedule and tile that win for
 * one L can lose for a mix.  No counterpart in the reference (the batching
 * extension). */
int wv_index_autotune_budgets(wv_index *idx, const float *d_queries, uint32_t Q, uint64_t seed,
                              const uint32_t *d_L_per_q, const uint32_t *d_k_per_q, uint32_t k_stride,
                              const wv_search_ext *ext, wv_search_config *best, double *best_ms);

/* hipGraph-captured search for serving loops: the launch of
 * wv_index_search_ext_dev (random entries, with the stop model of `ext` when
 * given; no traces) planned, checked and captured once with the index's
 * current tuning, then replayed on the context's stream with the same device
 * buffers.  The caller may rewrite the query rows (and valid per-query L/k,
 * within the L and k_stride checked here) between launches; the kernel
 * clamps per-query L and k to what was planned.  Creation runs the search
 * once eagerly.  launch: sync != 0 waits for the replay.  No counterpart in
 * the reference (a serving-loop convenience). */
typedef struct wv_search_graph wv_search_graph;
int wv_search_graph_create(wv_index *idx, const float *d_queries, uint32_t Q, uint64_t seed, uint32_t L,
                           uint32_t k, const uint32_t *d_L_per_q, const uint32_t *d_k_per_q,
                           uint32_t k_stride, uint32_t *d_out_ids, float *d_out_dist, uint32_t *d_evals,
                           uint32_t *d_hops, const wv_search_ext *ext, wv_search_graph **out);
int wv_search_graph_launch(wv_search_graph *g, int sync);
int wv_search_graph_destroy(wv_search_graph *g);

/* ---------------------------------------------------------------- MLP */
/* X: n x 8 raw features g (standardised with the model's mean/inv_std
 * inside), y: n labels in {0,1}.  Host arrays. */
int wv_mlp_grad(wv_ctx *ctx, const wv_term_model *m, const float *X, const float *y, uint32_t n,
                float pos_weight, float *grad /* 321: w1, b1, w2, b2 */, double *loss);
int wv_mlp_train(wv_ctx *ctx, const float *X, const float *y, uint32_t n,
                 const wv_train_params *tp, wv_term_model *model /* in: init; out */,
                 wv_train_stats *stats);
/* Logits exactly as the fused search step computes them (BF16 products,
 * FP32 accumulation). */
int wv_mlp_predict(wv_ctx *ctx, const wv_term_model *m, const float *X, uint32_t n, float *logits);

/* ---------------------------------------------------------------- one-shot */
/* The two stage entry points as single calls on host arrays. */
int wv_nndescent_build(wv_ctx *ctx, const float *base, uint32_t N, uint32_t d,
                       const wv_nnd_params *p, const uint32_t *init_ids, uint32_t *out_ids,
                       float *out_dist, uint32_t *out_deg, wv_build_stats *stats);
int wv_search_greedy(wv_ctx *ctx, const wv_graph *g, const float *base, uint32_t d,
                     const float *queries, uint32_t Q, const uint32_t *entries, uint64_t seed,
                     uint32_t L, uint32_t k, const uint32_t *L_per_q, const uint32_t *k_per_q,
                     uint32_t k_stride, uint32_t *out_ids, float *out_dist, uint32_t *evals,
                     uint32_t *hops, wv_search_stats *stats);

/* ---------------------------------------------------------------- oracle */
/* Exact kNN in FP64 on the device.  self_offset >= 0: query q is base row
 * q + self_offset and is skipped (kNN graph); -1: nothing skipped.  k <= 128. */
int wv_exact_knn(wv_ctx *ctx, const float *base, uint32_t N, const float *queries, uint32_t Q,
                 uint32_t d, uint32_t k, int64_t self_offset, uint32_t *out_ids,
                 double *out_dist, double *kernel_ms);

/* ---------------------------------------------------------------- measurement */
/* The rate this GPU sustains for the search's distance phase alone: random
 * row gathers scored against a query with the search's 64-row gather tile.
 * No counterpart in the reference; it gives the search's roofline an
 * achievable ceiling for its access pattern beside the HBM peak.
 * waves_per_cu wavefronts (1..16) are held resident on every CU (one a
 * workgroup, residency fixed through LDS), each drawing `batches` x 64
 * uniformly random ids of an N x d table (d <= 256, N >= 1) in device
 * memory.  *gbps = rows x 4d bytes / kernel time; *kernel_ms the time. */
int wv_gather_rate(wv_ctx *ctx, uint32_t N, uint32_t d, uint32_t waves_per_cu, uint32_t batches,
                   double *gbps, double *kernel_ms);

/* The rate this GPU sustains for loads of random lines: 16-byte loads at
 * uniformly random (hashed) positions of a device array of `bytes` bytes
 * (a multiple of 16, at least 1 MiB), 4 independent loads a thread per
 * round, `iters` rounds, 8 wavefronts resident a SIMD.  No counterpart in
 * the reference; it is the ceiling of the NN-Descent join, whose threshold
 * records, pool-id probes and candidate-slot updates are such loads.
 * *gaccess = loads / kernel time in G/s; *kernel_ms the time. */
int wv_line_rate(wv_ctx *ctx, uint64_t bytes, uint32_t iters, double *gaccess, double *kernel_ms);

#ifdef __cplusplus
}
#endif
#endif /* WEAVESS_HIP_H */
