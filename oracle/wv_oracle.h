/*
 * wv_oracle.h -- CPU restatement of WEAVESS's kNN-graph hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path in
 * weavess_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in the product links or calls it.
 *
 * Every function follows a reference routine, cited as path:line relative to
 * the WEAVESS source tree.  Documented deviations (shared with the GPU path):
 *   D1  ties are broken by id: every ordering is on (dist, id), never dist only
 *       (reference orders by dist only, include/weavess/index.h:74-80).
 *   D2  random numbers come from a counter-based generator (splitmix64
 *       finaliser on seed/stream/counter) instead of glibc rand()/mt19937
 *       (src/component_init.cpp:18, src/component_search_entry.cpp:15), so
 *       that a GPU wavefront can reproduce any draw without sequential state.
 *   D3  reverse samples above the cap R are chosen as the R smallest hash
 *       priorities instead of rand()%R overwrites in arrival order
 *       (src/component_refine.cpp:190-208), which makes the choice
 *       independent of thread order.
 *   D4  "build order" distance: see wvo_dist_canon.
 *   D5  update()'s reverse test reads each neighbour's sorted worst.  In the
 *       reference the loop that samples node n ends with make_heap of n's
 *       pool (src/component_refine.cpp:212), so a node visited later reads
 *       a heap leaf from n's pool.back() (:185, :196), not n's worst; with
 *       OpenMP threads which neighbours were visited first depends on the
 *       schedule.  wvo_nnd_params.sweep = 1 restates the single-thread
 *       order (neighbours with a smaller id read the leaf).
 *
 * Parity status of every exported function is listed in the table at the top
 * of wv_oracle.cpp.
 */
#ifndef WV_ORACLE_H
#define WV_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Distances ------------------------------------------------------------ */
/* Reference order: include/weavess/distance.h:12-36. */
float wvo_dist_ref(const float *a, const float *b, uint32_t d);
/* Build order shared with the GPU kernels. */
float wvo_dist_canon(const float *a, const float *b, uint32_t d);
/* FP64 squared L2 (exact-kNN oracle). */
double wvo_dist_f64(const float *a, const float *b, uint32_t d);

/* Random numbers -------------------------------------------------------- */
uint64_t wvo_stream(uint64_t seed, uint64_t tag, uint64_t idx);
uint32_t wvo_draw(uint64_t stream, uint64_t k);
/* GenRandom algorithm (include/weavess/util.h:13-27) on a counter stream. */
int wvo_gen_random(uint64_t stream, uint32_t *addr, uint32_t size, uint32_t N);
/* Same algorithm fed explicit draws (size+1 of them), for pinning against
 * the reference's own GenRandom driven by std::mt19937. */
int wvo_gen_random_draws(const uint32_t *draws, uint32_t *addr, uint32_t size, uint32_t N);

/* NN-Descent ------------------------------------------------------------ */
typedef struct {
    uint32_t K, L, S, R, iters;
    uint32_t cand_slots;   /* 0: exact bounded insertion (reference);      */
                           /* >0: power-of-two lossy slot table per node   */
    uint32_t sweep;        /* 0: every reverse test reads the neighbour's  */
                           /* sorted worst (two-phase, D5); 1: the         */
                           /* reference's single-thread sweep, in which a  */
                           /* neighbour already visited was re-heaped and  */
                           /* its pool.back() is a heap leaf               */
    uint32_t reserved;
    uint64_t seed;
} wvo_nnd_params;

typedef struct {
    uint64_t pairs;        /* distance evaluations in the joins            */
    uint64_t inserts;      /* accepted list updates (new entries kept)     */
} wvo_nnd_iter_stats;

/* Random S-NN initialisation (src/component_init.cpp:11-45).
 * ids/dist are N*S row-major, cnt[i] = S or S-1 (self skipped). */
int wvo_init_random(const float *base, uint32_t N, uint32_t d, uint32_t S,
                    uint64_t seed, uint32_t *ids, float *dist, uint32_t *cnt,
                    int nthreads);

/* InitRandom + RefineNNDescent + PruneNaive.
 * init_ids/init_dist/init_cnt: optional N*S initial graph (NULL: run
 * wvo_init_random with p->seed).  Outputs N*K ids/dist and deg[N].
 * iter_stats (optional) has p->iters entries. */
int wvo_nndescent(const float *base, uint32_t N, uint32_t d,
                  const wvo_nnd_params *p,
                  const uint32_t *init_ids, const float *init_dist,
                  const uint32_t *init_cnt,
                  uint32_t *out_ids, float *out_dist, uint32_t *out_deg,
                  wvo_nnd_iter_stats *iter_stats, int nthreads);

/* The state of wvo_nndescent after p->iters (join, update) rounds, before
 * the final sort and PruneNaive: each node's pool (N*L ids, dist and "new"
 * flags, pool_cnt[N] valid, ascending), M[N] (the sample window of the last
 * update, src/component_refine.cpp:166-174), and the sample lists that the
 * next join would read: nn_new (N*new_cap, new_cnt[N]) and nn_old
 * (N*old_cap, old_cnt[N]), forward samples in pool order followed by the
 * reverse samples in ascending id.  Returns -2 when a list exceeds its cap. */
int wvo_nndescent_state(const float *base, uint32_t N, uint32_t d,
                        const wvo_nnd_params *p,
                        const uint32_t *init_ids, const float *init_dist,
                        const uint32_t *init_cnt,
                        uint32_t *pool_ids, float *pool_dist, uint8_t *pool_flag,
                        uint32_t *pool_cnt, uint32_t *M,
                        uint32_t *nn_new, uint32_t *new_cnt, uint32_t new_cap,
                        uint32_t *nn_old, uint32_t *old_cnt, uint32_t old_cap,
                        wvo_nnd_iter_stats *iter_stats, int nthreads);

/* ComponentRefineEFANNA + PruneNaive (src/component_refine.cpp:959-1183):
 * pools from an initial graph (init_ids rows of init_stride <= L ids,
 * init_cnt[i] valid, or NULL for wvo_init_random's graph), sorted with self
 * and repeated ids dropped; the first join runs over 2S GenRandom samples
 * per node (stream tag "EFNA").  Needs 2S < N.  Outputs as wvo_nndescent. */
int wvo_nndescent_efanna(const float *base, uint32_t N, uint32_t d,
                         const wvo_nnd_params *p,
                         const uint32_t *init_ids, const uint32_t *init_cnt,
                         uint32_t init_stride,
                         uint32_t *out_ids, float *out_dist, uint32_t *out_deg,
                         wvo_nnd_iter_stats *iter_stats, int nthreads);

/* NSG refine (src/component_refine.cpp:393-540 with ComponentCandidateNSG,
 * ComponentPruneNSG and ComponentConnNSGDFS) of the kNN graph knn
 * (N*kstride ids, kdeg[N]) from navigating node ep: L = "L_refine",
 * R = "R_refine", C = "C_refine".  Writes N*wcap ids/dist and deg[N];
 * returns -2 with *width set when the widest row exceeds wcap. */
int wvo_nsg(const float *base, uint32_t N, uint32_t d, const uint32_t *knn, const uint32_t *kdeg,
            uint32_t kstride, uint32_t L, uint32_t R, uint32_t C, uint32_t ep, uint64_t seed,
            uint32_t *out_ids, float *out_dist, uint32_t *out_deg, uint32_t wcap, uint32_t *width,
            uint32_t *conn_edges, int nthreads);
/* One pass of ComponentRefineVAMANA (src/component_refine.cpp:794-871) on
 * graph ids/deg: the NSG candidate search, Vamana's prune (alpha, filled
 * to R with the nearest occluded candidates, component_prune.cpp:245-297),
 * batch reverse insertion with the same prune; no connectivity pass. */
int wvo_vamana(const float *base, uint32_t N, uint32_t d, const uint32_t *ids, const uint32_t *deg,
               uint32_t stride, uint32_t L, uint32_t R, uint32_t C, float alpha, uint32_t ep, uint64_t seed,
               uint32_t *out_ids, float *out_dist, uint32_t *out_deg, uint32_t wcap, uint32_t *width,
               int nthreads);
/* ComponentRefineDPG + ComponentConnReverse (src/component_refine.cpp:
 * 728-790, component_conn.cpp:228-268) on graph ids/deg with list length
 * range ("K" / 2 in the reference). */
int wvo_dpg(const float *base, uint32_t N, uint32_t d, const uint32_t *ids, const uint32_t *deg, uint32_t stride,
            uint32_t range, uint32_t *out_ids, float *out_dist, uint32_t *out_deg, uint32_t wcap, uint32_t *width,
            int nthreads);
/* ComponentRefineSSG (src/component_refine.cpp:550-727): Propagation2
 * candidates (L = "L_refine"), PruneSSG's angle rule (angle "A" degrees),
 * batch InterInsert, ComponentConnSSGDFS from n_try roots (written to eps).
 * wvo_ssg_threshold: the cosine threshold used. */
float wvo_ssg_threshold(float angle);
int wvo_ssg(const float *base, uint32_t N, uint32_t d, const uint32_t *knn, const uint32_t *kdeg, uint32_t kstride,
            uint32_t L, uint32_t R, float angle, uint32_t n_try, uint64_t seed, uint32_t *out_ids, float *out_dist,
            uint32_t *out_deg, uint32_t wcap, uint32_t *width, uint32_t *eps, uint32_t *conn_edges, int nthreads);
/* The connectivity pass alone, on graph ids/deg. */
int wvo_nsg_conn(const float *base, uint32_t N, uint32_t d, const uint32_t *ids, const uint32_t *deg,
                 uint32_t stride, uint32_t L, uint32_t C, uint32_t ep, uint64_t seed, uint32_t *out_ids,
                 float *out_dist, uint32_t *out_deg, uint32_t wcap, uint32_t *width, uint32_t *conn_edges,
                 int nthreads);

/* Greedy search --------------------------------------------------------- */
/* SearchEntryRand + RouteGreedy (src/component_search_entry.cpp:9-30,
 * src/component_route.cpp:9-53).  graph is N*deg_max ids, deg[N] valid
 * counts.  entries: Q*L ids or NULL (generated from seed).  L_per_q/k_per_q
 * may be NULL (then L/k are used for every query).  out_ids/out_dist are
 * Q*k_stride.  evals/hops may be NULL. */
int wvo_search_greedy(const uint32_t *graph, const uint32_t *deg,
                      uint32_t deg_max, const float *base, uint32_t N,
                      uint32_t d, const float *queries, uint32_t Q,
                      const uint32_t *entries, uint64_t seed,
                      uint32_t L, uint32_t k,
                      const uint32_t *L_per_q, const uint32_t *k_per_q,
                      uint32_t k_stride,
                      uint32_t *out_ids, float *out_dist,
                      uint32_t *evals, uint32_t *hops, int nthreads);

/* Learned termination (same layout as wv_term_model) --------------------- */
typedef struct {
    float w1[32][8];
    float b1[32];
    float w2[32];
    float b2;
    float mean[8];
    float inv_std[8];
    float logit_threshold;
    uint32_t min_hops;
} wvo_term_model;

/* Search-state features g, FP32 (restates the device formula). */
void wvo_term_features(float hops, float kth, float best, float entry, float kth4, float stale,
                       float L, float k, float *f8);
/* Gradient (321: w1, b1, w2, b2) and mean weighted BCE in FP64. */
int wvo_mlp_grad(const wvo_term_model *m, const float *X, const float *y, uint32_t n,
                 float pos_weight, double *grad, double *loss);
/* Logits: bf16 != 0 follows the device's BF16-product / FP32-sum order,
 * bf16 == 0 is plain FP64. */
int wvo_mlp_logit(const wvo_term_model *m, const float *X, uint32_t n, int bf16, double *out);

/* Greedy search with per-hop traces and/or the learned stop rule, the rule
 * evaluated after every 4th expansion.  model, trace (Q*cap*8), trace_len,
 * settle may be NULL. */
int wvo_search_greedy_ext(const uint32_t *graph, const uint32_t *deg, uint32_t deg_max,
                          const float *base, uint32_t N, uint32_t d, const float *queries,
                          uint32_t Q, const uint32_t *entries, uint64_t seed, uint32_t L,
                          uint32_t k, const uint32_t *L_per_q, const uint32_t *k_per_q,
                          uint32_t k_stride, uint32_t *out_ids, float *out_dist,
                          uint32_t *evals, uint32_t *hops, const wvo_term_model *model,
                          float *trace, uint32_t *trace_len, uint32_t *settle, uint32_t cap,
                          int nthreads);

/* Exact kNN ------------------------------------------------------------- */
/* FP64 brute force, ties by id.  self_offset >= 0: query q is base row
 * q+self_offset and is excluded (kNN graph); -1: no exclusion. */
int wvo_exact_knn(const float *base, uint32_t N, const float *queries,
                  uint32_t Q, uint32_t d, uint32_t k, int64_t self_offset,
                  uint32_t *out_ids, double *out_dist, int nthreads);

/* Metrics --------------------------------------------------------------- */
/* recall@k, src/builder.cpp:465-480. */
double wvo_recall(const uint32_t *res, uint32_t Q, uint32_t k,
                  uint32_t res_stride, const uint32_t *gt, uint32_t gt_dim);
/* graph quality, src/builder.cpp:760-776 (self already excluded in gt). */
double wvo_graph_quality(const uint32_t *ids, const uint32_t *deg,
                         uint32_t deg_stride, const uint32_t *rows,
                         uint32_t nrows, const uint32_t *gt, uint32_t gt_dim,
                         uint32_t control);

#ifdef __cplusplus
}
#endif
#endif
