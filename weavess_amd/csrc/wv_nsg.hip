// wv_nsg.hip -- NSG refine of a kNN graph on the GPU (ComponentRefineNSG,
// src/component_refine.cpp:393-540, with ComponentCandidateNSG,
// ComponentPruneNSG and ComponentConnNSGDFS).
//
//   1. candidates: per node, a greedy search for the node's own vector over
//      the kNN graph from ep's neighbours plus random ids, pool L, keeping
//      the C nearest distinct nodes it evaluated (the search kernel's
//      collection mode);
//   2. prune: those plus the node's kNN neighbours, sorted, first C; the
//      occlusion rule keeps up to R;
//   3. reverse insertion as one batch: every node's list grows by the nodes
//      that kept it, and a list above R is re-pruned;
//   4. connectivity: nodes unreachable from ep are linked to a reached node
//      near them (host traversal, GPU searches).
//
// Distances are the build-order FP32 distance everywhere, so every decision
// matches the CPU restatement bit for bit.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "wv_device.hpp"
#include "wv_internal.hpp"
#include "wv_nnd.hpp"
#include "wv_pool.hpp"

namespace wv {
namespace {

constexpr uint64_t TAG_NSGC = 0x4E534743ull;  // "NSGC": candidate-search entries
constexpr uint64_t TAG_NSGR = 0x4E534752ull;  // "NSGR": connectivity fallback root
constexpr uint64_t TAG_SSGE = 0x53534745ull;  // "SSGE": SSG's n_try roots
constexpr uint32_t NSG_TCAP = 1024;           // candidates of one reverse-insert re-prune
constexpr int NWPB = 1;                       // wavefronts per block: one node each, a finished one frees its slot at once

#define NSG_HIP(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) {                                                              \
            char m_[256];                                                                    \
            snprintf(m_, sizeof m_, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                     __LINE__);                                                              \
            return index_set_error(ix, e_ == hipErrorOutOfMemory ? WV_ENOMEM : WV_EHIP, m_); \
        }                                                                                    \
    } while (0)

struct NsgArgs {
    const float *base;
    uint32_t N, dp, nch;
    const uint32_t *knn;   // N x kstride kNN graph
    const uint32_t *kdeg;
    uint32_t kstride;
    uint32_t R, C;
    const uint64_t *full;  // batch x C collected keys
    uint32_t n0, nb;       // batch: nodes n0 .. n0+nb-1
    uint32_t *pr_ids;      // N x R pruned lists
    float *pr_dist;
    uint32_t *pr_cnt;
    const uint32_t *inoff, *insrc;  // reverse edges (CSR)
    const float *indist;
    uint32_t *out_ids;     // N x R after the reverse insertion
    float *out_dist;
    uint32_t *out_cnt;
    uint32_t lds_wave;
    float alpha;           // occlusion factor (NSG: 1)
    uint32_t add_knn;      // candidates include the node's kNN neighbours (NSG)
    uint32_t fill;         // top lists up with the nearest occluded candidates (Vamana)
    uint32_t angle;        // SSG's angle rule instead of occlusion
    float thr;             // its cosine threshold
    uint32_t L, HS, CAPC;  // SSG candidates: two-hop count, visited slots, candidate capacity
};

template <int CPL>
constexpr int rows_g() { return CPL >= 16 ? 1 : CPL >= 8 ? 2 : 4; }
constexpr int DPG_NQ = 4;  // DPG's prune: earlier entries per pass over the later ones

template <int CPL>
__device__ inline void load_row(const float *base, uint32_t dp, uint32_t nch, uint32_t id, float4 (&qv)[CPL]) {
    const float4 *qp = reinterpret_cast<const float4 *>(base + (size_t)id * dp);
    const int lane = lane_id();
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
        const int ch = (lane & 15) + 16 * c;
        qv[c] = ch < (int)nch ? qp[ch] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// Candidate states of the occlusion pass.
constexpr uint8_t ST_DROP = 0, ST_LIVE = 1, ST_KEPT = 2, ST_SELF = 3;

// Occlusion (component_prune.cpp:40-58, and Vamana's :245-297) over the
// sorted distinct keys cand[0..m) of one node (key distance = distance to
// the node), with st[i] the candidates' states (ST_LIVE or ST_SELF): the
// first live key is kept, and every kept key drops the later live keys p
// with alpha * d(kept, p) < d(node, p); at most R kept, written to res.
// With `fill`, a list shorter than R is topped up with the nearest dropped
// keys, in order.  Returns the list length.
// With `angle` (PruneSSG, component_prune.cpp:84-105) the test is instead
// cos(p, kept) > thr, the cosine from the law of cosines on squared
// distances: (d(p) + d(kept) - d(kept, p)) / 2 / sqrt(d(p) d(kept)).
template <int CPL>
__device__ int occlude(const float *base, uint32_t dp, uint32_t nch, const uint64_t *cand, int m, uint8_t *st,
                       int R, float alpha, bool fill, uint32_t *cid, uint32_t *cpos, uint64_t *ckey, uint64_t *res,
                       bool angle = false, float thr = 0.f) {
    const int lane = lane_id();
    int cnt = 0, from = 0;
    while (cnt < R) {
        int nxt = m;
        for (int t0 = from & ~(WAVE - 1); t0 < m; t0 += WAVE) {
            const int p = t0 + lane;
            const uint64_t b = __ballot(p >= from && p < m && st[p] == ST_LIVE);
            if (b) {
                nxt = t0 + __ffsll((unsigned long long)b) - 1;
                break;
            }
        }
        if (nxt >= m) break;
        const uint64_t ki = cand[nxt];
        if (lane == 0) {
            res[cnt] = ki;
            st[nxt] = ST_KEPT;
        }
        ++cnt;
        from = nxt + 1;
        if (cnt >= R) break;
        float4 qv[CPL];
        load_row<CPL>(base, dp, nch, key_id(ki), qv);
        for (int t0 = from & ~(WAVE - 1); t0 < m; t0 += WAVE) {
            const int p = t0 + lane;
            const bool ok = p >= from && p < m && st[p] == ST_LIVE;
            const uint64_t b = __ballot(ok);
            if (!b) continue;
            const int mm = __popcll(b);
            if (ok) {
                cid[lanes_below(b)] = key_id(cand[p]);
                cpos[lanes_below(b)] = (uint32_t)p;
            }
            wave_sync();
            rows_dist<CPL, rows_g<CPL>()>(base, dp, nch, qv, cid, mm, ckey);
            wave_sync();
            if (lane < mm) {
                const uint32_t pp = cpos[lane];
                const float dpn = key_dist(cand[pp]), djk = key_dist(ckey[lane]);
                bool drop;
                if (angle) {
                    const float dr = key_dist(ki);
                    drop = ((dpn + dr) - djk) / 2.0f / sqrtf(dpn * dr) > thr;
                } else {
                    drop = alpha * djk < dpn;
                }
                if (drop) st[pp] = ST_DROP;
            }
            wave_sync();
        }
    }
    wave_sync();
    if (fill)
        for (int t0 = 0; t0 < m && cnt < R; t0 += WAVE) {
            const int p = t0 + lane;
            const bool ok = p < m && st[p] == ST_DROP;
            const uint64_t b = __ballot(ok);
            const int pos = cnt + (int)lanes_below(b);
            if (ok && pos < R) res[pos] = cand[p];
            cnt = min(R, cnt + __popcll(b));
        }
    wave_sync();
    return cnt;
}

// The occlusion rule of `occlude` evaluated block by block: candidates in
// blocks of 64; at a block's start every key kept so far passes once over
// the block's live candidates, then the block's candidates are taken in
// order, each kept one passing over the rest of the block.  A candidate is
// kept or dropped by the same keys as in `occlude`, so the lists are equal;
// the work stops at the block where the R-th key is kept.  That pays when
// the rule drops few candidates (Vamana's alpha > 1) or when the R keys
// come early (the reverse re-prunes); SSG's angle rule keeps `occlude`.
template <int CPL>
__device__ int occlude_blocked(const float *base, uint32_t dp, uint32_t nch, const uint64_t *cand, int m,
                               uint8_t *st, int R, float alpha, bool fill, uint32_t *cid, uint32_t *cpos,
                               uint64_t *ckey, uint64_t *res) {
    const int lane = lane_id();
    int cnt = 0;
    // drop the block's live candidates that kept key k occludes
    auto pass = [&](uint64_t k, int from, int bend) {
        float4 qv[CPL];
        load_row<CPL>(base, dp, nch, key_id(k), qv);
        const int p = from + lane;
        const bool ok = p < bend && st[p] == ST_LIVE;
        const uint64_t b = __ballot(ok);
        if (!b) return;
        const int mm = __popcll(b);
        if (ok) {
            cid[lanes_below(b)] = key_id(cand[p]);
            cpos[lanes_below(b)] = (uint32_t)p;
        }
        wave_sync();
        rows_dist<CPL, rows_g<CPL>()>(base, dp, nch, qv, cid, mm, ckey);
        wave_sync();
        if (lane < mm) {
            const uint32_t pp = cpos[lane];
            if (alpha * key_dist(ckey[lane]) < key_dist(cand[pp])) st[pp] = ST_DROP;
        }
        wave_sync();
    };
    for (int b0 = 0; b0 < m && cnt < R; b0 += WAVE) {
        const int bend = min(m, b0 + WAVE);
        for (int r = 0; r < cnt; ++r) pass(res[r], b0, bend);
        int from = b0;
        while (cnt < R) {
            const uint64_t b = __ballot(from + lane < bend && st[from + lane] == ST_LIVE);
            if (!b) break;
            const int nxt = from + __ffsll((unsigned long long)b) - 1;
            const uint64_t ki = cand[nxt];
            if (lane == 0) {
                res[cnt] = ki;
                st[nxt] = ST_KEPT;
            }
            wave_sync();
            ++cnt;
            from = nxt + 1;
            if (cnt < R && from < bend) pass(ki, from, bend);
        }
    }
    wave_sync();
    if (fill)
        for (int t0 = 0; t0 < m && cnt < R; t0 += WAVE) {
            const int p = t0 + lane;
            const bool ok = p < m && st[p] == ST_DROP;
            const uint64_t b = __ballot(ok);
            const int pos = cnt + (int)lanes_below(b);
            if (ok && pos < R) res[pos] = cand[p];
            cnt = min(R, cnt + __popcll(b));
        }
    wave_sync();
    return cnt;
}

// Entries of the candidate searches of nodes[q] (q-th query; nodes null:
// node n0 + q): the h head ids (ep's list), then the ids of GenRandom(L) on
// the node's "NSGC" stream that are not in the head (sorted copy hs), in
// order, up to L.
__global__ void __launch_bounds__(256) nsg_entries_kernel(const uint32_t *nodes, uint32_t n0, uint32_t Q,
                                                          const uint32_t *head, const uint32_t *hs, uint32_t h,
                                                          uint32_t L, uint32_t N, uint64_t seed, uint32_t lds_wave,
                                                          uint32_t *out) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int wid = threadIdx.x >> 6, lane = lane_id();
    const uint32_t q = blockIdx.x * (blockDim.x >> 6) + wid;
    if (q >= Q) return;
    const uint32_t n = nodes ? nodes[q] : n0 + q;
    uint64_t *a = reinterpret_cast<uint64_t *>(smem + (size_t)wid * lds_wave);
    uint64_t *tmp = a + ((L + 63) & ~63u);
    gen_random_lds(stream_of(seed, TAG_NSGC, n), L, N, a, tmp);
    uint32_t *o = out + (size_t)q * L;
    for (uint32_t i = lane; i < h; i += WAVE) o[i] = head[i];
    uint32_t w = h;
    for (uint32_t c0 = 0; c0 < L && w < L; c0 += WAVE) {
        const uint32_t i = c0 + lane;
        const uint32_t x = i < L ? (uint32_t)a[i] : ID_NONE;
        bool keep = i < L;
        if (keep) {  // not in the head
            uint32_t lo = 0, hi = h;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (hs[mid] < x) lo = mid + 1;
                else hi = mid;
            }
            keep = !(lo < h && hs[lo] == x);
        }
        const uint64_t bm = __ballot(keep);
        const uint32_t pos = w + lanes_below(bm);
        if (keep && pos < L) o[pos] = x;
        w += (uint32_t)__popcll(bm);
    }
}

// PruneNSG (component_prune.cpp:20-66) for nodes n0 .. n0+nb-1: the
// collected keys plus (NSG) the node's kNN neighbours, distinct, sorted,
// first C; self skipped; occlusion to R (Vamana: with its alpha and fill).
template <int CPL>
__global__ void __launch_bounds__(256) nsg_prune_kernel(NsgArgs a) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int wid = threadIdx.x >> 6, lane = lane_id();
    const uint32_t q = blockIdx.x * (blockDim.x >> 6) + wid;
    if (q >= a.nb) return;
    const uint32_t n = a.n0 + q;
    uint8_t *my = smem + (size_t)wid * a.lds_wave;
    uint64_t *cand = reinterpret_cast<uint64_t *>(my);  // C
    uint64_t *ckey = cand + a.C;                         // 64
    uint64_t *tmp = ckey + WAVE;                         // 64
    uint64_t *res = tmp + WAVE;                          // R
    uint32_t *cid = reinterpret_cast<uint32_t *>(res + a.R);  // 64
    uint32_t *cpos = cid + WAVE;                              // 64
    uint8_t *alive = reinterpret_cast<uint8_t *>(cpos + WAVE);  // C
    const int C = (int)a.C;
    int len = 0;
    for (int t0 = 0; t0 < C; t0 += WAVE) {
        const int i = t0 + lane;
        const uint64_t k = i < C ? a.full[(size_t)q * a.C + i] : KEY_EMPTY;
        if (i < C) cand[i] = k;
        len += __popcll(__ballot(k != KEY_EMPTY));
    }
    wave_sync();
    float4 qv[CPL];
    load_row<CPL>(a.base, a.dp, a.nch, n, qv);
    const uint32_t dg = a.add_knn ? min(a.kdeg[n], a.kstride) : 0u;
    for (uint32_t c0 = 0; c0 < dg; c0 += WAVE) {
        const uint32_t id = c0 + lane < dg ? a.knn[(size_t)n * a.kstride + c0 + lane] : ID_NONE;
        const bool ok = id < a.N;
        const uint64_t b = __ballot(ok);
        const int m = __popcll(b);
        if (ok) cid[lanes_below(b)] = id;
        wave_sync();
        if (m == 0) continue;
        rows_dist<CPL, rows_g<CPL>()>(a.base, a.dp, a.nch, qv, cid, m, ckey);
        wave_sync();
        len = merge_distinct(cand, len, C, ckey, m, tmp);
    }
    for (int i = lane; i < len; i += WAVE) alive[i] = key_id(cand[i]) != n ? ST_LIVE : ST_SELF;
    wave_sync();
    // blocked when the rule drops few (alpha > 1); at alpha = 1 over a
    // random initial graph's candidates the push form measured faster
    const int cnt = a.alpha > 1.0f
                        ? occlude_blocked<CPL>(a.base, a.dp, a.nch, cand, len, alive, (int)a.R, a.alpha, a.fill != 0,
                                               cid, cpos, ckey, res)
                        : occlude<CPL>(a.base, a.dp, a.nch, cand, len, alive, (int)a.R, a.alpha, a.fill != 0, cid,
                                       cpos, ckey, res);
    for (int i = lane; i < cnt; i += WAVE) {
        a.pr_ids[(size_t)n * a.R + i] = key_id(res[i]);
        a.pr_dist[(size_t)n * a.R + i] = key_dist(res[i]);
    }
    if (lane == 0) a.pr_cnt[n] = (uint32_t)cnt;
}

__global__ void nsg_rev_count_kernel(const uint32_t *pr_ids, const uint32_t *pr_cnt, uint32_t N, uint32_t R,
                                     uint32_t *incnt) {
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x)
        for (uint32_t j = 0; j < pr_cnt[n]; ++j) atomicAdd(&incnt[pr_ids[(size_t)n * R + j]], 1u);
}

__global__ void nsg_rev_scatter_kernel(const uint32_t *pr_ids, const float *pr_dist, const uint32_t *pr_cnt,
                                       uint32_t N, uint32_t R, const uint32_t *inoff, uint32_t *incur,
                                       uint32_t *insrc, float *indist) {
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x)
        for (uint32_t j = 0; j < pr_cnt[n]; ++j) {
            const uint32_t t = pr_ids[(size_t)n * R + j];
            const uint32_t pos = inoff[t] + atomicAdd(&incur[t], 1u);
            insrc[pos] = n;
            indist[pos] = pr_dist[(size_t)n * R + j];
        }
}

// InterInsert (component_refine.cpp:474-540) as one batch: node t's list
// grows by every s whose pruned list holds t and that t's list does not
// hold, in ascending s; above R, own + incoming are sorted, cut to the
// NSG_TCAP nearest and re-pruned by occlusion.
template <int CPL>
__global__ void __launch_bounds__(256) nsg_insert_kernel(NsgArgs a) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int wid = threadIdx.x >> 6, lane = lane_id();
    const uint32_t t = blockIdx.x * (blockDim.x >> 6) + wid;
    if (t >= a.N) return;
    uint8_t *my = smem + (size_t)wid * a.lds_wave;
    uint64_t *cand = reinterpret_cast<uint64_t *>(my);  // NSG_TCAP
    uint64_t *ckey = cand + NSG_TCAP;                    // 64
    uint64_t *tmp = ckey + WAVE;                         // 64
    uint64_t *res = tmp + WAVE;                          // R
    const uint32_t ow = a.C >> 1;                        // power of two >= R
    uint64_t *own = res + a.R;                           // ow
    uint32_t *cid = reinterpret_cast<uint32_t *>(own + ow);   // 64
    uint32_t *cpos = cid + WAVE;                              // 64
    uint32_t *srt = cpos + WAVE;                              // pow2 >= 2R
    uint8_t *alive = reinterpret_cast<uint8_t *>(srt + a.C);  // NSG_TCAP (a.C holds the sort width here)
    const int R = (int)a.R;
    const int no = (int)a.pr_cnt[t];
    // the node's list as sorted keys (a Vamana list, topped up with dropped
    // candidates, is not in key order): duplicate tests and the re-prune's
    // merge need the order
    for (uint32_t i = lane; i < ow; i += WAVE)
        own[i] = (int)i < no ? make_key(a.pr_dist[(size_t)t * a.R + i], a.pr_ids[(size_t)t * a.R + i]) : KEY_EMPTY;
    wave_sync();
    lds_sort64(own, (int)ow);
    const uint32_t s0 = a.inoff[t], s1 = a.inoff[t + 1];
    // incoming edges not already in the list
    int nnew = 0;
    for (uint32_t c0 = s0; c0 < s1; c0 += WAVE) {
        const uint32_t i = c0 + lane;
        bool fresh = false;
        if (i < s1) {
            const uint64_t k = make_key(a.indist[i], a.insrc[i]);
            const int p = lower_bound_e(own, no, k);
            fresh = !(p < no && own[p] == k);
        }
        nnew += __popcll(__ballot(fresh));
    }
    if (no + nnew <= R) {
        // own list, then the new sources in ascending id order
        const uint32_t w = a.C;  // power of two >= 2R, >= the segment here
        for (uint32_t i = lane; i < w; i += WAVE) {
            uint32_t v = ID_NONE;
            if (s0 + i < s1) {
                const uint64_t k = make_key(a.indist[s0 + i], a.insrc[s0 + i]);
                const int p = lower_bound_e(own, no, k);
                if (!(p < no && own[p] == k)) v = a.insrc[s0 + i];
            }
            srt[i] = v;
        }
        wave_sync();
        lds_sort32(srt, (int)w);
        for (int i = lane; i < no; i += WAVE) {  // the list in its own order
            a.out_ids[(size_t)t * a.R + i] = a.pr_ids[(size_t)t * a.R + i];
            a.out_dist[(size_t)t * a.R + i] = a.pr_dist[(size_t)t * a.R + i];
        }
        for (int i = lane; i < nnew; i += WAVE) {
            const uint32_t s = srt[i];
            // the distance of edge s -> t: find it in the segment
            float dd = 0.f;
            for (uint32_t j = s0; j < s1; ++j)
                if (a.insrc[j] == s) {
                    dd = a.indist[j];
                    break;
                }
            a.out_ids[(size_t)t * a.R + no + i] = s;
            a.out_dist[(size_t)t * a.R + no + i] = dd;
        }
        if (lane == 0) a.out_cnt[t] = (uint32_t)(no + nnew);
        return;
    }
    // re-prune: the NSG_TCAP smallest distinct keys of own + incoming
    int len = 0;
    for (int i = lane; i < no; i += WAVE) cand[i] = own[i];
    len = min(no, (int)NSG_TCAP);
    wave_sync();
    for (uint32_t c0 = s0; c0 < s1; c0 += WAVE) {
        const uint32_t i = c0 + lane;
        if (i < s1) ckey[lane] = make_key(a.indist[i], a.insrc[i]);
        wave_sync();
        len = merge_distinct(cand, len, (int)NSG_TCAP, ckey, (int)min((uint32_t)WAVE, s1 - c0), tmp);
    }
    for (int i = lane; i < len; i += WAVE) alive[i] = ST_LIVE;
    wave_sync();
    const int cnt = !a.angle
                        ? occlude_blocked<CPL>(a.base, a.dp, a.nch, cand, len, alive, R, a.alpha, a.fill != 0, cid,
                                               cpos, ckey, res)
                        : occlude<CPL>(a.base, a.dp, a.nch, cand, len, alive, R, a.alpha, a.fill != 0, cid, cpos,
                                       ckey, res, a.angle != 0, a.thr);
    for (int i = lane; i < cnt; i += WAVE) {
        a.out_ids[(size_t)t * a.R + i] = key_id(res[i]);
        a.out_dist[(size_t)t * a.R + i] = key_dist(res[i]);
    }
    if (lane == 0) a.out_cnt[t] = (uint32_t)cnt;
}

// The build-order distance key of rows u and v (one wavefront).
template <int CPL>
__global__ void pair_key_kernel(const float *base, uint32_t dp, uint32_t nch, uint32_t u, uint32_t v, uint64_t *out) {
    __shared__ uint32_t cid[WAVE];
    __shared__ uint64_t ck[WAVE];
    float4 qv[CPL];
    load_row<CPL>(base, dp, nch, u, qv);
    if (lane_id() == 0) cid[0] = v;
    wave_sync();
    rows_dist<CPL, rows_g<CPL>()>(base, dp, nch, qv, cid, 1, ck);
    wave_sync();
    if (lane_id() == 0) out[0] = ck[0];
}

hipError_t pair_key(int cpl, const float *base, uint32_t dp, uint32_t nch, uint32_t u, uint32_t v, uint64_t *out,
                    hipStream_t s) {
    switch (cpl) {
    case 1: hipLaunchKernelGGL(pair_key_kernel<1>, dim3(1), dim3(WAVE), 0, s, base, dp, nch, u, v, out); break;
    case 2: hipLaunchKernelGGL(pair_key_kernel<2>, dim3(1), dim3(WAVE), 0, s, base, dp, nch, u, v, out); break;
    case 4: hipLaunchKernelGGL(pair_key_kernel<4>, dim3(1), dim3(WAVE), 0, s, base, dp, nch, u, v, out); break;
    case 8: hipLaunchKernelGGL(pair_key_kernel<8>, dim3(1), dim3(WAVE), 0, s, base, dp, nch, u, v, out); break;
    default: hipLaunchKernelGGL(pair_key_kernel<16>, dim3(1), dim3(WAVE), 0, s, base, dp, nch, u, v, out); break;
    }
    return hipGetLastError();
}

__global__ void gather_rows_kernel(const float *base, uint32_t dp, const uint32_t *ids, uint32_t n, float *out) {
    const size_t total = (size_t)n * dp;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x)
        out[i] = base[(size_t)ids[i / dp] * dp + i % dp];
}

template <class K>
hipError_t set_lds_k(K k, size_t bytes) {
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes);
}

struct DBuf {
    std::vector<void *> ps;
    ~DBuf() {
        for (void *p : ps) (void)hipFree(p);
    }
    template <class T>
    hipError_t get(T **p, size_t n) {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess) {
            ps.push_back(q);
            *p = static_cast<T *>(q);
        }
        return e;
    }
};

uint32_t pow2_ge(uint32_t v, uint32_t lo) {
    uint32_t p = lo;
    while (p < v) p <<= 1;
    return p;
}

template <int CPL>
hipError_t launch_prune(const NsgArgs &a, hipStream_t s) {
    const size_t lds = (size_t)NWPB * a.lds_wave;
    hipError_t e = set_lds_k(nsg_prune_kernel<CPL>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(nsg_prune_kernel<CPL>, dim3((a.nb + NWPB - 1) / NWPB), dim3(64 * NWPB), lds, s, a);
    return hipGetLastError();
}
template <int CPL>
hipError_t launch_insert(const NsgArgs &a, hipStream_t s) {
    const size_t lds = (size_t)NWPB * a.lds_wave;
    hipError_t e = set_lds_k(nsg_insert_kernel<CPL>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(nsg_insert_kernel<CPL>, dim3((a.N + NWPB - 1) / NWPB), dim3(64 * NWPB), lds, s, a);
    return hipGetLastError();
}

hipError_t dispatch_prune(int cpl, const NsgArgs &a, hipStream_t s) {
    switch (cpl) {
    case 1: return launch_prune<1>(a, s);
    case 2: return launch_prune<2>(a, s);
    case 4: return launch_prune<4>(a, s);
    case 8: return launch_prune<8>(a, s);
    default: return launch_prune<16>(a, s);
    }
}
hipError_t dispatch_insert(int cpl, const NsgArgs &a, hipStream_t s) {
    switch (cpl) {
    case 1: return launch_insert<1>(a, s);
    case 2: return launch_insert<2>(a, s);
    case 4: return launch_insert<4>(a, s);
    case 8: return launch_insert<8>(a, s);
    default: return launch_insert<16>(a, s);
    }
}

// ep's first min(L, degree) distinct in-range ids.
std::vector<uint32_t> head_of(const uint32_t *row, uint32_t deg, uint32_t L, uint32_t N) {
    std::vector<uint32_t> h;
    for (uint32_t j = 0; j < deg && h.size() < L; ++j)
        if (row[j] < N && std::find(h.begin(), h.end(), row[j]) == h.end()) h.push_back(row[j]);
    return h;
}

double ms_since(std::chrono::high_resolution_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0).count();
}

// SSG's candidates and prune for node n (Propagation2,
// component_candidate.cpp:89-108, then PruneSSG, component_prune.cpp:71-122):
// the neighbours' neighbours in list order, each once and never n, until
// L; then n's own list; distinct, sorted; the angle rule keeps up to R.
// The visited ids sit in an exact LDS hash (at most L + 1 ids in 2(L + 1)+
// slots, so it never fills).
template <int CPL>
__global__ void __launch_bounds__(256) ssg_prune_kernel(NsgArgs a) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int wid = threadIdx.x >> 6, lane = lane_id();
    const uint32_t n = blockIdx.x * (blockDim.x >> 6) + wid;
    if (n >= a.N) return;
    uint8_t *my = smem + (size_t)wid * a.lds_wave;
    uint64_t *cand = reinterpret_cast<uint64_t *>(my);  // CAPC
    uint64_t *ckey = cand + a.CAPC;                      // 64
    uint64_t *tmp = ckey + WAVE;                         // 64
    uint64_t *res = tmp + WAVE;                          // R
    uint32_t *cid = reinterpret_cast<uint32_t *>(res + a.R);  // 64
    uint32_t *cpos = cid + WAVE;                              // 64
    uint32_t *vis = cpos + WAVE;                              // HS
    uint8_t *st = reinterpret_cast<uint8_t *>(vis + a.HS);    // CAPC
    const uint32_t hmask = a.HS - 1;
    for (uint32_t i = lane; i < a.HS; i += WAVE) vis[i] = ID_NONE;
    wave_sync();
    if (lane == 0) {
        uint32_t h = fmix32(n) & hmask;
        vis[h] = n;
    }
    wave_sync();
    float4 qv[CPL];
    load_row<CPL>(a.base, a.dp, a.nch, n, qv);
    int len = 0;
    uint32_t taken = 0;
    const uint32_t dn = min(a.kdeg[n], a.kstride);
    for (uint32_t i = 0; i < dn && taken < a.L; ++i) {
        const uint32_t nid = a.knn[(size_t)n * a.kstride + i];
        if (nid >= a.N) continue;
        const uint32_t d2 = min(a.kdeg[nid], a.kstride);
        for (uint32_t c0 = 0; c0 < d2 && taken < a.L; c0 += WAVE) {
            const uint32_t v = c0 + lane < d2 ? a.knn[(size_t)nid * a.kstride + c0 + lane] : ID_NONE;
            bool fresh = v < a.N;
            if (fresh) {  // not yet visited (lookup)
                uint32_t h = fmix32(v) & hmask;
                for (;;) {
                    const uint32_t e = vis[h];
                    if (e == v) {
                        fresh = false;
                        break;
                    }
                    if (e == ID_NONE) break;
                    h = (h + 1) & hmask;
                }
            }
            const uint64_t b = __ballot(fresh);
            const uint32_t room = a.L - taken;
            const bool take = fresh && lanes_below(b) < room;
            const uint64_t tb = __ballot(take);
            const int m = __popcll(tb);
            if (take) {
                cid[lanes_below(tb)] = v;
                uint32_t h = fmix32(v) & hmask;
                while (atomicCAS(&vis[h], ID_NONE, v) != ID_NONE) h = (h + 1) & hmask;
            }
            wave_sync();
            if (m == 0) continue;
            rows_dist<CPL, rows_g<CPL>()>(a.base, a.dp, a.nch, qv, cid, m, ckey);
            wave_sync();
            len = merge_distinct(cand, len, (int)a.CAPC, ckey, m, tmp);
            taken += (uint32_t)m;
        }
    }
    for (uint32_t c0 = 0; c0 < dn; c0 += WAVE) {  // the node's own list
        const uint32_t id = c0 + lane < dn ? a.knn[(size_t)n * a.kstride + c0 + lane] : ID_NONE;
        const bool ok = id < a.N && id != n;
        const uint64_t b = __ballot(ok);
        const int m = __popcll(b);
        if (ok) cid[lanes_below(b)] = id;
        wave_sync();
        if (m == 0) continue;
        rows_dist<CPL, rows_g<CPL>()>(a.base, a.dp, a.nch, qv, cid, m, ckey);
        wave_sync();
        len = merge_distinct(cand, len, (int)a.CAPC, ckey, m, tmp);
    }
    for (int i = lane; i < len; i += WAVE) st[i] = ST_LIVE;
    wave_sync();
    const int cnt = occlude<CPL>(a.base, a.dp, a.nch, cand, len, st, (int)a.R, 1.0f, false, cid, cpos, ckey, res,
                                 true, a.thr);
    for (int i = lane; i < cnt; i += WAVE) {
        a.pr_ids[(size_t)n * a.R + i] = key_id(res[i]);
        a.pr_dist[(size_t)n * a.R + i] = key_dist(res[i]);
    }
    if (lane == 0) a.pr_cnt[n] = (uint32_t)cnt;
}

// Distance keys of node pairs (u[i], v[i]): one wavefront a pair.
template <int CPL>
__global__ void __launch_bounds__(256) pair_keys_kernel(const float *base, uint32_t dp, uint32_t nch,
                                                        const uint32_t *u, const uint32_t *v, uint32_t P,
                                                        uint64_t *out) {
    __shared__ uint32_t cid[4][WAVE];
    __shared__ uint64_t ck[4][WAVE];
    const int wid = threadIdx.x >> 6;
    const uint32_t i = blockIdx.x * (blockDim.x >> 6) + wid;
    if (i >= P) return;
    float4 qv[CPL];
    load_row<CPL>(base, dp, nch, u[i], qv);
    if (lane_id() == 0) cid[wid][0] = v[i];
    wave_sync();
    rows_dist<CPL, rows_g<CPL>()>(base, dp, nch, qv, cid[wid], 1, ck[wid]);
    wave_sync();
    if (lane_id() == 0) out[i] = ck[wid][0];
}

// One level of a breadth-first closure over the pruned graph (gid rows of
// gcnt ids) plus link lists (CSR xoff / xid): one wavefront per frontier
// node; each out-neighbour not yet reached is marked and appended to the
// next frontier (one counter atomic per wavefront batch).
__global__ void __launch_bounds__(256) ssg_bfs_kernel(const uint32_t *gid, const uint32_t *gcnt, uint32_t R,
                                                      const uint32_t *xoff, const uint32_t *xid, uint32_t *flag,
                                                      const uint32_t *fin, uint32_t nin, uint32_t *fout,
                                                      uint32_t *nout) {
    const int wid = threadIdx.x >> 6, lane = lane_id();
    const uint32_t i = blockIdx.x * (blockDim.x >> 6) + wid;
    if (i >= nin) return;
    const uint32_t u = fin[i];
    const uint32_t cu = gcnt[u], x0 = xoff[u], cx = xoff[u + 1] - x0;
    for (uint32_t j0 = 0; j0 < cu + cx; j0 += WAVE) {
        const uint32_t j = j0 + (uint32_t)lane;
        uint32_t w = ID_NONE;
        if (j < cu) w = gid[(size_t)u * R + j];
        else if (j < cu + cx) w = xid[x0 + (j - cu)];
        const bool fresh = w != ID_NONE && flag[w] == 0u && atomicExch(&flag[w], 1u) == 0u;
        const uint64_t b = __ballot(fresh);
        if (b == 0) continue;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(nout, (uint32_t)__popcll(b));
        base = (uint32_t)__shfl((int)base, 0);
        if (fresh) fout[base + lanes_below(b)] = w;
    }
}

// Labels of the nodes a closure left unreached: label[w] = the smallest
// unreached id with a path to w (reached nodes hold N and are skipped; no
// path from a reached node leads back to an unreached one).  One relaxation
// pass over the unreached nodes' out-edges, a wavefront per node; *changed
// is set when a label drops.  Repeated until nothing changes.
__global__ void __launch_bounds__(256) ssg_label_kernel(const uint32_t *gid, const uint32_t *gcnt, uint32_t R,
                                                        const uint32_t *xoff, const uint32_t *xid,
                                                        const uint32_t *flag, uint32_t *label, uint32_t N,
                                                        uint32_t *changed) {
    const int wid = threadIdx.x >> 6, lane = lane_id();
    const uint32_t u = blockIdx.x * (blockDim.x >> 6) + wid;
    if (u >= N || flag[u]) return;
    const uint32_t lu = label[u];
    const uint32_t cu = gcnt[u], x0 = xoff[u], cx = xoff[u + 1] - x0;
    bool dropped = false;
    for (uint32_t j = (uint32_t)lane; j < cu + cx; j += WAVE) {
        const uint32_t w = j < cu ? gid[(size_t)u * R + j] : xid[x0 + (j - cu)];
        if (!flag[w] && lu < label[w]) {
            atomicMin(&label[w], lu);
            dropped = true;
        }
    }
    if (__ballot(dropped) && lane == 0) atomicOr(changed, 1u);
}

// label[x] = flag[x] ? N : x
__global__ void ssg_label_init_kernel(const uint32_t *flag, uint32_t *label, uint32_t N) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < N) label[x] = flag[x] ? N : x;
}

// Closure of `root` over rows gid (gcnt ids of R slots) plus link lists
// (CSR xoff / xid): reached flags (u32) into d_flag, a breadth-first level
// per launch.  d_f0 / d_f1 hold N ids each, d_n one counter.
hipError_t closure_dev(const uint32_t *gid, const uint32_t *gcnt, uint32_t R, const uint32_t *d_xoff,
                       const uint32_t *d_xid, uint32_t N, uint32_t root, uint32_t *d_flag, uint32_t *d_f0,
                       uint32_t *d_f1, uint32_t *d_n, hipStream_t s) {
    hipError_t e = hipMemsetAsync(d_flag, 0, (size_t)N * 4, s);
    const uint32_t one = 1u;
    if (e == hipSuccess) e = hipMemcpyAsync(d_flag + root, &one, 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_f0, &root, 4, hipMemcpyHostToDevice, s);
    uint32_t nin = 1;
    uint32_t *fin = d_f0, *fout = d_f1;
    while (e == hipSuccess && nin > 0) {
        e = hipMemsetAsync(d_n, 0, 4, s);
        if (e != hipSuccess) break;
        hipLaunchKernelGGL(ssg_bfs_kernel, dim3((nin + 3) / 4), dim3(256), 0, s, gid, gcnt, R, d_xoff, d_xid, d_flag,
                           fin, nin, fout, d_n);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(&nin, d_n, 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        std::swap(fin, fout);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);  // `one` and `root` outlive their copies
    return e;
}

template <int CPL>
hipError_t launch_ssg(const NsgArgs &a, hipStream_t s) {
    const size_t lds = (size_t)NWPB * a.lds_wave;
    hipError_t e = set_lds_k(ssg_prune_kernel<CPL>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(ssg_prune_kernel<CPL>, dim3((a.N + NWPB - 1) / NWPB), dim3(64 * NWPB), lds, s, a);
    return hipGetLastError();
}

hipError_t pair_keys(int cpl, const float *base, uint32_t dp, uint32_t nch, const uint32_t *u, const uint32_t *v,
                     uint32_t P, uint64_t *out, hipStream_t s) {
    const dim3 g((P + 3) / 4), b(256);
    switch (cpl) {
    case 1: hipLaunchKernelGGL(pair_keys_kernel<1>, g, b, 0, s, base, dp, nch, u, v, P, out); break;
    case 2: hipLaunchKernelGGL(pair_keys_kernel<2>, g, b, 0, s, base, dp, nch, u, v, P, out); break;
    case 4: hipLaunchKernelGGL(pair_keys_kernel<4>, g, b, 0, s, base, dp, nch, u, v, P, out); break;
    case 8: hipLaunchKernelGGL(pair_keys_kernel<8>, g, b, 0, s, base, dp, nch, u, v, P, out); break;
    default: hipLaunchKernelGGL(pair_keys_kernel<16>, g, b, 0, s, base, dp, nch, u, v, P, out); break;
    }
    return hipGetLastError();
}

// DPG's prune (component_prune.cpp:124-186) of node n's list: the first
// len = min(size, 2 range) distinct keys (self dropped); hit[j] counts the
// earlier entries i with d(p_i, p_j) < d(n, p_j); cut is the (range+1)-th
// smallest hit; the entries with hit <= cut are kept in order, at most
// range (all of them when len <= range).
// rows_dist for NQ query rows at once: each gathered row serves all of
// them.  dout[q * 64 + r] is row r's distance to query q, computed with
// rows_dist's operations in rows_dist's order.
template <int CPL, int G, int NQ>
__device__ inline void rows_dist_multi(const float *__restrict__ base, uint32_t dp, uint32_t nch,
                                       const float4 (&qv)[NQ][CPL], const uint32_t *ids, int m, float *dout) {
    const int lane = lane_id(), sub = lane & 15, grp = lane >> 4;
    for (int r0 = 0; r0 < m; r0 += 4 * G) {
        float4 v[G][CPL];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int row = r0 + 4 * g + grp;
            const uint32_t id = row < m ? ids[row] : 0u;
            const float4 *p = reinterpret_cast<const float4 *>(base + (size_t)id * dp);
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                const int ch = sub + 16 * c;
                v[g][c] = (row < m && ch < (int)nch) ? p[ch] : make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int row = r0 + 4 * g + grp;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                float acc = 0.f;
#pragma unroll
                for (int c = 0; c < CPL; ++c)
                    if (sub + 16 * c < (int)nch) acc = acc + chunk_sq(qv[q][c], v[g][c]);
                acc = fold16(acc);
                if (sub == 0 && row < m) dout[q * WAVE + row] = acc;
            }
        }
    }
}

template <int CPL>
__global__ void __launch_bounds__(256) dpg_prune_kernel(NsgArgs a) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int wid = threadIdx.x >> 6, lane = lane_id();
    const uint32_t n = blockIdx.x * (blockDim.x >> 6) + wid;
    if (n >= a.N) return;
    const int R = (int)a.R, cap = 2 * R;
    const uint32_t sp = a.C;  // power of two >= 2R: sort width of the hits
    uint8_t *my = smem + (size_t)wid * a.lds_wave;
    uint64_t *p = reinterpret_cast<uint64_t *>(my);  // 2R
    uint64_t *ckey = p + cap;                         // 64
    uint64_t *tmp = ckey + WAVE;                      // 64
    uint32_t *cid = reinterpret_cast<uint32_t *>(tmp + WAVE);  // 64
    uint32_t *hit = cid + WAVE;                                // 2R
    uint32_t *srt = hit + cap;                                 // sp
    float4 qv[CPL];
    load_row<CPL>(a.base, a.dp, a.nch, n, qv);
    int len = 0;
    const uint32_t dg = min(a.kdeg[n], a.kstride);
    for (uint32_t c0 = 0; c0 < dg; c0 += WAVE) {
        const uint32_t id = c0 + lane < dg ? a.knn[(size_t)n * a.kstride + c0 + lane] : ID_NONE;
        const bool ok = id < a.N && id != n;
        const uint64_t b = __ballot(ok);
        const int m = __popcll(b);
        if (ok) cid[lanes_below(b)] = id;
        wave_sync();
        if (m == 0) continue;
        rows_dist<CPL, rows_g<CPL>()>(a.base, a.dp, a.nch, qv, cid, m, ckey);
        wave_sync();
        len = merge_distinct(p, len, cap, ckey, m, tmp);
    }
    int cnt = 0;
    if (len <= R) {
        for (int i = lane; i < len; i += WAVE) {
            a.pr_ids[(size_t)n * a.R + i] = key_id(p[i]);
            a.pr_dist[(size_t)n * a.R + i] = key_dist(p[i]);
        }
        cnt = len;
    } else {
        for (int i = lane; i < len; i += WAVE) hit[i] = 0;
        wave_sync();
        // entries i .. i + DPG_NQ - 1 against the later ones at once: each
        // later row is gathered once for the group (ckey and tmp, 1 KB,
        // hold the group's distances)
        float *dout = reinterpret_cast<float *>(ckey);
        for (int i0 = 0; i0 + 1 < len; i0 += DPG_NQ) {
            float4 pv[DPG_NQ][CPL];
#pragma unroll
            for (int k = 0; k < DPG_NQ; ++k) load_row<CPL>(a.base, a.dp, a.nch, key_id(p[min(i0 + k, len - 1)]), pv[k]);
            for (int t0 = i0 + 1; t0 < len; t0 += WAVE) {
                const int m = min(WAVE, len - t0);
                if (lane < m) cid[lane] = key_id(p[t0 + lane]);
                wave_sync();
                rows_dist_multi<CPL, rows_g<CPL>(), DPG_NQ>(a.base, a.dp, a.nch, pv, cid, m, dout);
                wave_sync();
                if (lane < m) {
                    const int j = t0 + lane;
                    const float dn = key_dist(p[j]);
                    uint32_t h = 0;
#pragma unroll
                    for (int k = 0; k < DPG_NQ; ++k)  // pairs i < j, i <= len - 2
                        if (i0 + k < j && i0 + k + 1 < len && dout[k * WAVE + lane] < dn) ++h;
                    hit[j] += h;
                }
                wave_sync();
            }
        }
        for (uint32_t i = lane; i < sp; i += WAVE) srt[i] = (int)i < len ? hit[i] : 0xFFFFFFFFu;
        wave_sync();
        lds_sort32(srt, (int)sp);
        const uint32_t cut = srt[R];
        for (int t0 = 0; t0 < len && cnt < R; t0 += WAVE) {
            const int j = t0 + lane;
            const bool keep = j < len && hit[j] <= cut;
            const uint64_t b = __ballot(keep);
            const int pos = cnt + (int)lanes_below(b);
            if (keep && pos < R) {
                a.pr_ids[(size_t)n * a.R + pos] = key_id(p[j]);
                a.pr_dist[(size_t)n * a.R + pos] = key_dist(p[j]);
            }
            cnt = min(R, cnt + __popcll(b));
        }
    }
    if (lane == 0) a.pr_cnt[n] = (uint32_t)cnt;
}

// ConnReverse (component_conn.cpp:228-268): node t's list plus every
// reverse edge, as keys in t's CSR segment; own entries first, then the
// reverse ones at a per-node cursor.
__global__ void dpg_union_fill_kernel(const uint32_t *pr_ids, const float *pr_dist, const uint32_t *pr_cnt, uint32_t N,
                                      uint32_t R, const uint32_t *off, uint32_t *cur, uint64_t *keys) {
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
        const uint32_t c = pr_cnt[n];
        for (uint32_t j = 0; j < c; ++j) {
            const uint32_t t = pr_ids[(size_t)n * R + j];
            const float dd = pr_dist[(size_t)n * R + j];
            keys[off[n] + j] = make_key(dd, t);
            keys[off[t] + pr_cnt[t] + atomicAdd(&cur[t], 1u)] = make_key(dd, n);
        }
    }
}

// Distinct keys of each sorted segment, one wavefront a node: count
// (write == 0) or write the padded row of width W.  Lanes take 64 keys at a
// time; a key is kept when it differs from the key before it in the segment
// (the first against KEY_EMPTY), so each kept key's column is the number of
// kept keys before it, and the row and its padding are stored coalesced.
__global__ void __launch_bounds__(256) dpg_union_rows_kernel(const uint64_t *keys, const uint32_t *off, uint32_t N,
                                                             uint32_t *ucnt, uint32_t W, uint32_t *gid, float *gdist,
                                                             uint32_t *gdeg, int write) {
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint32_t nw = (gridDim.x * blockDim.x) / WAVE;
    for (uint32_t t = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE; t < N; t += nw) {
        const uint32_t b = off[t], e = off[t + 1];
        uint32_t u = 0;
        for (uint32_t i0 = b; i0 < e; i0 += WAVE) {
            const uint32_t i = i0 + lane;
            const bool in = i < e;
            const uint64_t k = in ? keys[i] : KEY_EMPTY;
            const uint64_t kp = in && i > b ? keys[i - 1] : KEY_EMPTY;
            const bool keep = in && k != kp;
            const uint64_t bm = __ballot(keep);
            if (write && keep) {
                const size_t c = (size_t)t * W + u + lanes_below(bm);
                gid[c] = key_id(k);
                gdist[c] = key_dist(k);
            }
            u += (uint32_t)__popcll(bm);
        }
        if (write) {
            for (uint32_t i = u + lane; i < W; i += WAVE) {
                gid[(size_t)t * W + i] = ID_NONE;
                gdist[(size_t)t * W + i] = -1.0f;
            }
            if (lane == 0) gdeg[t] = u;
        } else if (lane == 0) {
            ucnt[t] = u;
        }
    }
}

template <int CPL>
hipError_t launch_dpg(const NsgArgs &a, hipStream_t s) {
    const size_t lds = (size_t)NWPB * a.lds_wave;
    hipError_t e = set_lds_k(dpg_prune_kernel<CPL>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(dpg_prune_kernel<CPL>, dim3((a.N + NWPB - 1) / NWPB), dim3(64 * NWPB), lds, s, a);
    return hipGetLastError();
}


// Reverse insertion of the pruned lists in `a` (pr_*) into a.out_* with the
// rule in `a` (NSG, Vamana or SSG re-prune); see nsg_insert_kernel.
int reverse_insert(wv_index *ix, NsgArgs &a, int cpl, DBuf &B, hipStream_t s) {
    const uint32_t N = a.N, R = a.R;
    uint32_t *incnt = nullptr, *inoff = nullptr, *incur = nullptr, *insrc = nullptr;
    float *indist = nullptr;
    NSG_HIP(B.get(&incnt, (size_t)N + 1));
    NSG_HIP(B.get(&inoff, (size_t)N + 1));
    NSG_HIP(B.get(&incur, N));
    NSG_HIP(B.get(&insrc, (size_t)N * R));
    NSG_HIP(B.get(&indist, (size_t)N * R));
    NSG_HIP(hipMemsetAsync(incnt, 0, ((size_t)N + 1) * 4, s));
    NSG_HIP(hipMemsetAsync(incur, 0, (size_t)N * 4, s));
    hipLaunchKernelGGL(nsg_rev_count_kernel, dim3(1024), dim3(256), 0, s, a.pr_ids, a.pr_cnt, N, R, incnt);
    NSG_HIP(hipGetLastError());
    size_t scan_bytes = 0;
    NSG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, incnt, inoff, N + 1, s));
    uint8_t *scan_tmp = nullptr;
    NSG_HIP(B.get(&scan_tmp, scan_bytes));
    NSG_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, incnt, inoff, N + 1, s));
    hipLaunchKernelGGL(nsg_rev_scatter_kernel, dim3(1024), dim3(256), 0, s, a.pr_ids, a.pr_dist, a.pr_cnt, N, R,
                       inoff, incur, insrc, indist);
    NSG_HIP(hipGetLastError());
    a.inoff = inoff;
    a.insrc = insrc;
    a.indist = indist;
    NSG_HIP(B.get(&a.out_ids, (size_t)N * R));
    NSG_HIP(B.get(&a.out_dist, (size_t)N * R));
    NSG_HIP(B.get(&a.out_cnt, N));
    const uint32_t sw = pow2_ge(2 * R, 64);  // sort width of new sources; own keys sort in sw / 2 >= R
    NsgArgs ai = a;
    ai.C = sw;
    ai.lds_wave = (uint32_t)((((size_t)NSG_TCAP * 8 + 64 * 8 * 2 + (size_t)R * 8 + (size_t)(sw / 2) * 8 +
                               64 * 4 * 2 + (size_t)sw * 4 + NSG_TCAP) + 15) & ~(size_t)15);
    NSG_HIP(dispatch_insert(cpl, ai, s));
    NSG_HIP(hipStreamSynchronize(s));
    return WV_OK;
}

// GenRandom (include/weavess/util.h:13-27) of `size` distinct ids below N on
// counter stream st, as the device and the CPU restatement draw it.
void gen_random_host(uint64_t st, uint32_t size, uint32_t N, uint32_t *out) {
    std::vector<uint32_t> v(size);
    for (uint32_t i = 0; i < size; ++i) v[i] = draw(st, i) % (N - size);
    std::sort(v.begin(), v.end());
    for (uint32_t i = 1; i < size; ++i)
        if (v[i] <= v[i - 1]) v[i] = v[i - 1] + 1;
    const uint32_t off = draw(st, size) % N;
    for (uint32_t i = 0; i < size; ++i) out[i] = (uint32_t)(((uint64_t)v[i] + off) % N);
}

}  // namespace

// The result graph's rows, a wavefront per node: the node's list slots
// (ids / dist, cnt of R), then ID_NONE / -1 up to W.
__global__ void __launch_bounds__(256) result_rows_kernel(const uint32_t *ids, const float *dist,
                                                          const uint32_t *cnt, uint32_t N, uint32_t R, uint32_t W,
                                                          uint32_t *gid, float *gdist) {
    const uint32_t i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (i >= N) return;
    const uint32_t c = cnt[i];
    for (uint32_t j = (uint32_t)lane_id(); j < W; j += WAVE) {
        const bool in = j < c;
        gid[(size_t)i * W + j] = in ? ids[(size_t)i * R + j] : ID_NONE;
        gdist[(size_t)i * W + j] = in ? dist[(size_t)i * R + j] : -1.0f;
    }
}

// Links after the lists: slot lslot[k] of node lnode[k] takes lid[k], ldist[k].
__global__ void result_links_kernel(const uint32_t *lnode, const uint32_t *lslot, const uint32_t *lid,
                                    const float *ldist, uint32_t nl, uint32_t W, uint32_t *gid, float *gdist) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nl) return;
    const size_t e = (size_t)lnode[k] * W + lslot[k];
    gid[e] = lid[k];
    gdist[e] = ldist[k];
}

// Links of a connectivity step, in the order they follow each node's list.
struct LinkSet {
    std::vector<uint32_t> node, slot, id;
    std::vector<float> dist;
};

// Store the refined graph in the index, assembled on the device from the
// lists in a (out_ids / out_dist / out_cnt) plus the links; deg is each
// node's final degree, W the widest.
int store_graph(wv_index *ix, const NsgArgs &a, uint32_t N, uint32_t R, uint32_t W, const std::vector<uint32_t> &deg,
                const LinkSet &links, DBuf &B, hipStream_t s) {
    uint32_t *gi = nullptr, *gg = nullptr;
    float *gd = nullptr;
    int rc = index_alloc_graph_dev(ix, W, &gi, &gd, &gg);  // frees the kNN graph
    if (rc) return rc;
    hipLaunchKernelGGL(result_rows_kernel, dim3((N + 3) / 4), dim3(256), 0, s, a.out_ids, a.out_dist, a.out_cnt, N, R,
                       W, gi, gd);
    NSG_HIP(hipGetLastError());
    const uint32_t nl = (uint32_t)links.node.size();
    if (nl) {
        uint32_t *d_node = nullptr, *d_slot = nullptr, *d_id = nullptr;
        float *d_dist = nullptr;
        NSG_HIP(B.get(&d_node, nl));
        NSG_HIP(B.get(&d_slot, nl));
        NSG_HIP(B.get(&d_id, nl));
        NSG_HIP(B.get(&d_dist, nl));
        NSG_HIP(hipMemcpyAsync(d_node, links.node.data(), (size_t)nl * 4, hipMemcpyHostToDevice, s));
        NSG_HIP(hipMemcpyAsync(d_slot, links.slot.data(), (size_t)nl * 4, hipMemcpyHostToDevice, s));
        NSG_HIP(hipMemcpyAsync(d_id, links.id.data(), (size_t)nl * 4, hipMemcpyHostToDevice, s));
        NSG_HIP(hipMemcpyAsync(d_dist, links.dist.data(), (size_t)nl * 4, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(result_links_kernel, dim3((nl + 255) / 256), dim3(256), 0, s, d_node, d_slot, d_id, d_dist,
                           nl, W, gi, gd);
        NSG_HIP(hipGetLastError());
    }
    NSG_HIP(hipMemcpyAsync(gg, deg.data(), (size_t)N * 4, hipMemcpyHostToDevice, s));
    NSG_HIP(hipStreamSynchronize(s));
    return WV_OK;
}

int nsg_build(wv_index *ix, const wv_nsg_params *p, wv_nsg_stats *stats, const uint32_t *knn, const uint32_t *kdeg,
              uint32_t kstride, const RefineRule &rule) {
    // knn / kdeg: the graph being refined, on the device (kstride ids a row)
    const IndexView v = index_view(ix);
    const hipStream_t s = v.stream;
    const uint32_t N = v.N, L = p->L, R = p->R, C = p->C, ep = p->ep;
    const uint32_t nch = v.dp / 4;
    const int cpl = nch <= 16 ? 1 : nch <= 32 ? 2 : nch <= 64 ? 4 : nch <= 128 ? 8 : 16;
    NSG_HIP(hipStreamSynchronize(s));
    const auto t_all = std::chrono::high_resolution_clock::now();
    DBuf B;
    NsgArgs a{};
    a.base = v.base;
    a.N = N;
    a.dp = v.dp;
    a.nch = nch;
    a.knn = knn;
    a.kdeg = kdeg;
    a.kstride = kstride;
    a.R = R;
    a.C = C;
    a.alpha = rule.alpha;
    a.add_knn = rule.add_knn ? 1u : 0u;
    a.fill = rule.fill ? 1u : 0u;
    NSG_HIP(B.get(&a.pr_ids, (size_t)N * R));
    NSG_HIP(B.get(&a.pr_dist, (size_t)N * R));
    NSG_HIP(B.get(&a.pr_cnt, N));

    // entries head: ep's kNN list
    std::vector<uint32_t> eprow(kstride), epdeg(1);
    NSG_HIP(hipMemcpy(eprow.data(), knn + (size_t)ep * kstride, (size_t)kstride * 4, hipMemcpyDeviceToHost));
    NSG_HIP(hipMemcpy(epdeg.data(), kdeg + ep, 4, hipMemcpyDeviceToHost));
    std::vector<uint32_t> head = head_of(eprow.data(), std::min(epdeg[0], kstride), L, N), hs;
    hs = head;
    std::sort(hs.begin(), hs.end());
    uint32_t *d_head = nullptr, *d_hs = nullptr;
    NSG_HIP(B.get(&d_head, head.size()));
    NSG_HIP(B.get(&d_hs, head.size()));
    if (!head.empty()) {
        NSG_HIP(hipMemcpy(d_head, head.data(), head.size() * 4, hipMemcpyHostToDevice));
        NSG_HIP(hipMemcpy(d_hs, hs.data(), hs.size() * 4, hipMemcpyHostToDevice));
    }
    const uint32_t ent_lds = (uint32_t)((((L + 63) & ~63u) + 64) * 8);
    NSG_HIP(set_lds_k(nsg_entries_kernel, (size_t)NWPB * ent_lds));

    // 1-2. candidates and prune, in node batches
    const uint32_t BN = std::min<uint32_t>(N, 1u << 18);
    uint32_t *d_ent = nullptr, *d_evals = nullptr;
    uint64_t *d_full = nullptr;
    NSG_HIP(B.get(&d_ent, (size_t)BN * L));
    NSG_HIP(B.get(&d_full, (size_t)BN * C));
    NSG_HIP(B.get(&d_evals, BN));
    a.full = d_full;
    a.lds_wave = (uint32_t)((((size_t)C * 8 + 64 * 8 * 2 + (size_t)R * 8 + 64 * 4 * 2 + C) + 15) & ~(size_t)15);
    uint64_t evals = 0;
    double search_ms = 0, prune_ms = 0;
    std::vector<uint32_t> hev(BN);
    for (uint32_t n0 = 0; n0 < N; n0 += BN) {
        const uint32_t nb = std::min(BN, N - n0);
        auto t0 = std::chrono::high_resolution_clock::now();
        hipLaunchKernelGGL(nsg_entries_kernel, dim3((nb + NWPB - 1) / NWPB), dim3(64 * NWPB), NWPB * ent_lds, s,
                           nullptr, n0, nb, d_head, d_hs, (uint32_t)head.size(), L, N, p->seed, ent_lds, d_ent);
        NSG_HIP(hipGetLastError());
        int rc = search_collect(ix, knn, kdeg, kstride, v.base + (size_t)n0 * v.dp, nb, d_ent, L, C, d_full,
                                d_evals);
        if (rc) return rc;
        NSG_HIP(hipMemcpyAsync(hev.data(), d_evals, (size_t)nb * 4, hipMemcpyDeviceToHost, s));
        NSG_HIP(hipStreamSynchronize(s));
        for (uint32_t i = 0; i < nb; ++i) evals += hev[i];
        search_ms += ms_since(t0);
        t0 = std::chrono::high_resolution_clock::now();
        a.n0 = n0;
        a.nb = nb;
        NSG_HIP(dispatch_prune(cpl, a, s));
        NSG_HIP(hipStreamSynchronize(s));
        prune_ms += ms_since(t0);
    }

    // 3. reverse insertion
    auto t0 = std::chrono::high_resolution_clock::now();
    {
        const int rc = reverse_insert(ix, a, cpl, B, s);
        if (rc) return rc;
    }
    const double insert_ms = ms_since(t0);

    // 4. connectivity (component_conn.cpp:14-98); the lists come to the host
    // only when links may be needed (their reach is marked there)
    t0 = std::chrono::high_resolution_clock::now();
    std::vector<uint32_t> gid(rule.conn ? (size_t)N * R : 0), gcnt(N);
    if (rule.conn) NSG_HIP(hipMemcpy(gid.data(), a.out_ids, gid.size() * 4, hipMemcpyDeviceToHost));
    NSG_HIP(hipMemcpy(gcnt.data(), a.out_cnt, (size_t)N * 4, hipMemcpyDeviceToHost));
    std::vector<std::vector<std::pair<uint32_t, float>>> extra(N);
    std::vector<uint8_t> flag(N, 0);
    std::vector<uint32_t> st;
    auto mark = [&](uint32_t root) {
        st.assign(1, root);
        flag[root] = 1;
        while (!st.empty()) {
            const uint32_t u = st.back();
            st.pop_back();
            for (uint32_t j = 0; j < gcnt[u]; ++j) {
                const uint32_t w = gid[(size_t)u * R + j];
                if (!flag[w]) {
                    flag[w] = 1;
                    st.push_back(w);
                }
            }
            for (const auto &e : extra[u])
                if (!flag[e.first]) {
                    flag[e.first] = 1;
                    st.push_back(e.first);
                }
        }
    };
    std::vector<uint32_t> un;
    if (rule.conn) {
        // what ep reaches, on the device (the later links' closures are small
        // and stay on the host)
        uint32_t *d_flag = nullptr, *d_f0 = nullptr, *d_f1 = nullptr, *d_n = nullptr, *d_xoff = nullptr;
        NSG_HIP(B.get(&d_flag, N));
        NSG_HIP(B.get(&d_f0, N));
        NSG_HIP(B.get(&d_f1, N));
        NSG_HIP(B.get(&d_n, 1));
        NSG_HIP(B.get(&d_xoff, (size_t)N + 1));
        NSG_HIP(hipMemsetAsync(d_xoff, 0, ((size_t)N + 1) * 4, s));
        NSG_HIP(closure_dev(a.out_ids, a.out_cnt, R, d_xoff, nullptr, N, ep, d_flag, d_f0, d_f1, d_n, s));
        std::vector<uint32_t> flag32(N);
        NSG_HIP(hipMemcpy(flag32.data(), d_flag, (size_t)N * 4, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < N; ++i) {
            flag[i] = flag32[i] != 0;
            if (!flag[i]) un.push_back(i);
        }
    }
    uint32_t conn_edges = 0;
    if (!un.empty()) {
        const uint32_t U = (uint32_t)un.size();
        std::vector<uint32_t> h2 = head_of(gid.data() + (size_t)ep * R, gcnt[ep], L, N), hs2(h2);
        std::sort(hs2.begin(), hs2.end());
        uint32_t *d_h2 = nullptr, *d_hs2 = nullptr, *d_un = nullptr, *d_ent2 = nullptr, *d_ev2 = nullptr;
        uint64_t *d_full2 = nullptr;
        float *d_q2 = nullptr;
        NSG_HIP(B.get(&d_h2, h2.size()));
        NSG_HIP(B.get(&d_hs2, h2.size()));
        NSG_HIP(B.get(&d_un, U));
        NSG_HIP(B.get(&d_ent2, (size_t)U * L));
        NSG_HIP(B.get(&d_full2, (size_t)U * C));
        NSG_HIP(B.get(&d_ev2, U));
        NSG_HIP(B.get(&d_q2, (size_t)U * v.dp));
        if (!h2.empty()) {
            NSG_HIP(hipMemcpy(d_h2, h2.data(), h2.size() * 4, hipMemcpyHostToDevice));
            NSG_HIP(hipMemcpy(d_hs2, hs2.data(), hs2.size() * 4, hipMemcpyHostToDevice));
        }
        NSG_HIP(hipMemcpy(d_un, un.data(), (size_t)U * 4, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(gather_rows_kernel, dim3(1024), dim3(256), 0, s, v.base, v.dp, d_un, U, d_q2);
        NSG_HIP(hipGetLastError());
        hipLaunchKernelGGL(nsg_entries_kernel, dim3((U + NWPB - 1) / NWPB), dim3(64 * NWPB), NWPB * ent_lds, s, d_un,
                           0u, U, d_h2, d_hs2, (uint32_t)h2.size(), L, N, p->seed, ent_lds, d_ent2);
        NSG_HIP(hipGetLastError());
        // search the pruned graph as it stands before any link
        int rc = search_collect(ix, a.out_ids, a.out_cnt, R, d_q2, U, d_ent2, L, C, d_full2, d_ev2);
        if (rc) return rc;
        std::vector<uint64_t> F((size_t)U * C);
        NSG_HIP(hipMemcpy(F.data(), d_full2, F.size() * 8, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < U; ++i) {
            const uint32_t u = un[i];
            if (flag[u]) continue;
            uint32_t root = N;
            float dd = 0.f;
            for (uint32_t j = 0; j < C; ++j) {
                const uint64_t k = F[(size_t)i * C + j];
                if (k == KEY_EMPTY) break;
                if (flag[key_id(k)]) {
                    root = key_id(k);
                    dd = key_dist(k);
                    break;
                }
            }
            if (root == N) {
                const uint64_t sx = stream_of(p->seed, TAG_NSGR, u);
                for (uint64_t tt = 0; tt < (1ull << 26) && root == N; ++tt) {
                    const uint32_t x = draw(sx, tt) % N;
                    if (flag[x]) root = x;
                }
                if (root == N) root = ep;
                // the edge's distance, the build-order way on the device
                uint64_t *d_k = nullptr;
                NSG_HIP(B.get(&d_k, 1));
                NSG_HIP(pair_key(cpl, v.base, v.dp, nch, root, u, d_k, s));
                NSG_HIP(hipStreamSynchronize(s));  // the context's stream does not wait for hipMemcpy
                uint64_t kk = 0;
                NSG_HIP(hipMemcpy(&kk, d_k, 8, hipMemcpyDeviceToHost));
                dd = key_dist(kk);
            }
            extra[root].push_back({u, dd});
            ++conn_edges;
            mark(u);
        }
    }
    // result graph: pruned + inserted lists, then the links
    uint32_t W = 1;
    std::vector<uint32_t> fdeg(N);
    LinkSet links;
    for (uint32_t i = 0; i < N; ++i) {
        fdeg[i] = gcnt[i] + (uint32_t)extra[i].size();
        W = std::max<uint32_t>(W, fdeg[i]);
        for (size_t j = 0; j < extra[i].size(); ++j) {
            links.node.push_back(i);
            links.slot.push_back(gcnt[i] + (uint32_t)j);
            links.id.push_back(extra[i][j].first);
            links.dist.push_back(extra[i][j].second);
        }
    }
    int rc = store_graph(ix, a, N, R, W, fdeg, links, B, s);
    if (rc) return rc;
    const double conn_ms = ms_since(t0);
    if (stats) {
        stats->total_ms = ms_since(t_all);
        stats->search_ms = search_ms;
        stats->prune_ms = prune_ms;
        stats->insert_ms = insert_ms;
        stats->conn_ms = conn_ms;
        stats->evals = evals;
        stats->conn_edges = conn_edges;
        stats->width = W;
    }
    return WV_OK;
}

int ssg_build(wv_index *ix, const wv_ssg_params *p, uint32_t *eps_out, wv_nsg_stats *stats, const uint32_t *knn,
              const uint32_t *kdeg, uint32_t kstride) {
    const IndexView v = index_view(ix);
    const hipStream_t s = v.stream;
    const uint32_t N = v.N, R = p->R, nch = v.dp / 4;
    const int cpl = nch <= 16 ? 1 : nch <= 32 ? 2 : nch <= 64 ? 4 : nch <= 128 ? 8 : 16;
    NSG_HIP(hipStreamSynchronize(s));
    const auto t_all = std::chrono::high_resolution_clock::now();
    DBuf B;
    NsgArgs a{};
    a.base = v.base;
    a.N = N;
    a.dp = v.dp;
    a.nch = nch;
    a.knn = knn;
    a.kdeg = kdeg;
    a.kstride = kstride;
    a.R = R;
    a.alpha = 1.0f;
    a.angle = 1;
    // component_refine.cpp:638-639: cos(A / 180 * pi), A in FP32, pi and the
    // cosine in FP64, rounded to FP32
    a.thr = (float)std::cos(p->angle / 180 * std::acos(-1.0));
    a.L = p->L;
    a.HS = pow2_ge(2 * (p->L + 1), 64);
    a.CAPC = p->L + kstride;
    a.lds_wave = (uint32_t)((((size_t)a.CAPC * 8 + 64 * 8 * 2 + (size_t)R * 8 + 64 * 4 * 2 + (size_t)a.HS * 4 +
                              a.CAPC) + 15) & ~(size_t)15);
    if ((size_t)NWPB * a.lds_wave > 160 * 1024)
        return index_set_error(ix, WV_ENOTSUP, "L_refine + kNN degree too large for LDS");
    NSG_HIP(B.get(&a.pr_ids, (size_t)N * R));
    NSG_HIP(B.get(&a.pr_dist, (size_t)N * R));
    NSG_HIP(B.get(&a.pr_cnt, N));
    switch (cpl) {
    case 1: NSG_HIP(launch_ssg<1>(a, s)); break;
    case 2: NSG_HIP(launch_ssg<2>(a, s)); break;
    case 4: NSG_HIP(launch_ssg<4>(a, s)); break;
    case 8: NSG_HIP(launch_ssg<8>(a, s)); break;
    default: NSG_HIP(launch_ssg<16>(a, s)); break;
    }
    NSG_HIP(hipStreamSynchronize(s));
    const double prune_ms = ms_since(t_all);
    auto t0 = std::chrono::high_resolution_clock::now();
    int rc = reverse_insert(ix, a, cpl, B, s);
    if (rc) return rc;
    const double insert_ms = ms_since(t0);

    // ComponentConnSSGDFS (component_conn.cpp:169-226), per root:
    //  1. the root's closure, on the device, a breadth-first level per launch
    //     (what a root reaches does not depend on the order of the visit);
    //  2. the reference then links the smallest unreached node `low` from
    //     the smallest reached node with room, reaches low's closure, and
    //     repeats.  A node x it leaves unreached is reached at the step of
    //     label(x), the smallest unreached id with a path to x, and low takes
    //     exactly the values v with label(v) == v: any smaller unreached
    //     node reaching v, or anything reaching it, would have been linked
    //     first and reached v.  Labels are propagated on the device;
    //  3. the host walks the labels in ascending order, linking and reaching
    //     as the reference does.
    t0 = std::chrono::high_resolution_clock::now();
    std::vector<uint32_t> gcnt(N);
    NSG_HIP(hipMemcpy(gcnt.data(), a.out_cnt, (size_t)N * 4, hipMemcpyDeviceToHost));
    std::vector<std::vector<uint32_t>> extra(N);
    std::vector<uint32_t> lu, lv;  // links lu[i] -> lv[i]
    std::vector<uint32_t> roots(p->n_try);
    gen_random_host(stream_of(p->seed, TAG_SSGE, 0), p->n_try, N, roots.data());
    uint32_t *d_flag = nullptr, *d_f0 = nullptr, *d_f1 = nullptr, *d_n = nullptr, *d_xoff = nullptr,
             *d_xid = nullptr, *d_label = nullptr;
    NSG_HIP(B.get(&d_flag, N));
    NSG_HIP(B.get(&d_f0, N));
    NSG_HIP(B.get(&d_f1, N));
    NSG_HIP(B.get(&d_n, 1));
    NSG_HIP(B.get(&d_xoff, (size_t)N + 1));
    NSG_HIP(B.get(&d_label, N));
    std::vector<uint32_t> xoff((size_t)N + 1), xid, flag32(N), label(N), boff((size_t)N + 1), bnode;
    size_t xcap = 0;
    std::vector<uint8_t> flag(N);
    std::vector<uint32_t> dgr(gcnt);  // current degree (links included)
    for (uint32_t r = 0; r < p->n_try; ++r) {
        // links so far, as CSR, for the device passes
        xid.clear();
        for (uint32_t i = 0; i < N; ++i) {
            xoff[i] = (uint32_t)xid.size();
            xid.insert(xid.end(), extra[i].begin(), extra[i].end());
        }
        xoff[N] = (uint32_t)xid.size();
        if (xid.size() > xcap) {
            xcap = std::max<size_t>(xid.size(), 2 * xcap);
            NSG_HIP(B.get(&d_xid, xcap));
        }
        NSG_HIP(hipMemcpyAsync(d_xoff, xoff.data(), xoff.size() * 4, hipMemcpyHostToDevice, s));
        if (!xid.empty()) NSG_HIP(hipMemcpyAsync(d_xid, xid.data(), xid.size() * 4, hipMemcpyHostToDevice, s));
        // 1. the root's closure
        NSG_HIP(closure_dev(a.out_ids, a.out_cnt, R, d_xoff, d_xid, N, roots[r], d_flag, d_f0, d_f1, d_n, s));
        // 2. labels of the unreached nodes (labels only drop, so this ends)
        hipLaunchKernelGGL(ssg_label_init_kernel, dim3((N + 255) / 256), dim3(256), 0, s, d_flag, d_label, N);
        NSG_HIP(hipGetLastError());
        for (uint32_t changed = 1; changed;) {
            NSG_HIP(hipMemsetAsync(d_n, 0, 4, s));
            hipLaunchKernelGGL(ssg_label_kernel, dim3((N + 3) / 4), dim3(256), 0, s, a.out_ids, a.out_cnt, R, d_xoff,
                               d_xid, d_flag, d_label, N, d_n);
            NSG_HIP(hipGetLastError());
            NSG_HIP(hipMemcpyAsync(&changed, d_n, 4, hipMemcpyDeviceToHost, s));
            NSG_HIP(hipStreamSynchronize(s));
        }
        NSG_HIP(hipMemcpyAsync(flag32.data(), d_flag, (size_t)N * 4, hipMemcpyDeviceToHost, s));
        NSG_HIP(hipMemcpyAsync(label.data(), d_label, (size_t)N * 4, hipMemcpyDeviceToHost, s));
        NSG_HIP(hipStreamSynchronize(s));
        // 3. unreached nodes bucketed by label, then the reference's walk
        std::fill(boff.begin(), boff.end(), 0u);
        uint32_t best = N;  // the smallest reached node with fewer than R neighbours
        for (uint32_t x = 0; x < N; ++x) {
            flag[x] = flag32[x] != 0;
            if (flag[x]) {
                if (dgr[x] < R && best == N) best = x;
            } else {
                ++boff[label[x] + 1];
            }
        }
        for (uint32_t v = 0; v < N; ++v) boff[v + 1] += boff[v];
        bnode.resize(boff[N]);
        {
            std::vector<uint32_t> cur(boff.begin(), boff.end() - 1);
            for (uint32_t x = 0; x < N; ++x)
                if (!flag[x]) bnode[cur[label[x]]++] = x;
        }
        for (uint32_t v = 0; v < N; ++v) {
            if (flag[v] || label[v] != v) continue;  // not a `low` of this root
            if (best < N) {
                const uint32_t j = best;
                extra[j].push_back(v);
                lu.push_back(j);
                lv.push_back(v);
                if (++dgr[j] >= R) {  // the next reached node with room: every one is above j
                    uint32_t b = j + 1;
                    while (b < N && !(flag[b] && dgr[b] < R)) ++b;
                    best = b;
                }
            }
            // v's closure among the unreached: the nodes labelled v
            for (uint32_t i = boff[v]; i < boff[v + 1]; ++i) {
                const uint32_t x = bnode[i];
                flag[x] = 1;
                if (dgr[x] < R && x < best) best = x;
            }
        }
    }
    // the links' distances, the build-order way on the device
    std::vector<uint64_t> lk(lu.size());
    if (!lu.empty()) {
        uint32_t *d_u = nullptr, *d_v = nullptr;
        uint64_t *d_k = nullptr;
        NSG_HIP(B.get(&d_u, lu.size()));
        NSG_HIP(B.get(&d_v, lv.size()));
        NSG_HIP(B.get(&d_k, lu.size()));
        NSG_HIP(hipMemcpy(d_u, lu.data(), lu.size() * 4, hipMemcpyHostToDevice));
        NSG_HIP(hipMemcpy(d_v, lv.data(), lv.size() * 4, hipMemcpyHostToDevice));
        NSG_HIP(pair_keys(cpl, v.base, v.dp, nch, d_u, d_v, (uint32_t)lu.size(), d_k, s));
        NSG_HIP(hipStreamSynchronize(s));
        NSG_HIP(hipMemcpy(lk.data(), d_k, lk.size() * 8, hipMemcpyDeviceToHost));
    }
    // result graph: pruned + inserted lists, then each node's links in the
    // order they were made
    std::vector<uint32_t> fdeg(gcnt);
    LinkSet links;
    for (size_t i = 0; i < lu.size(); ++i) {
        links.node.push_back(lu[i]);
        links.slot.push_back(fdeg[lu[i]]++);
        links.id.push_back(lv[i]);
        links.dist.push_back(key_dist(lk[i]));
    }
    uint32_t W = 1;
    for (uint32_t i = 0; i < N; ++i) W = std::max<uint32_t>(W, fdeg[i]);
    rc = store_graph(ix, a, N, R, W, fdeg, links, B, s);
    if (rc) return rc;
    const double conn_ms = ms_since(t0);
    if (eps_out) std::memcpy(eps_out, roots.data(), (size_t)p->n_try * 4);
    if (stats) {
        *stats = wv_nsg_stats{};
        stats->total_ms = ms_since(t_all);
        stats->prune_ms = prune_ms;
        stats->insert_ms = insert_ms;
        stats->conn_ms = conn_ms;
        stats->conn_edges = (uint32_t)lu.size();
        stats->width = W;
    }
    return WV_OK;
}

int dpg_build(wv_index *ix, uint32_t range, wv_nsg_stats *stats, const uint32_t *knn, const uint32_t *kdeg,
              uint32_t kstride) {
    const IndexView v = index_view(ix);
    const hipStream_t s = v.stream;
    const uint32_t N = v.N, R = range, nch = v.dp / 4;
    const int cpl = nch <= 16 ? 1 : nch <= 32 ? 2 : nch <= 64 ? 4 : nch <= 128 ? 8 : 16;
    NSG_HIP(hipStreamSynchronize(s));
    const auto t_all = std::chrono::high_resolution_clock::now();
    DBuf B;
    NsgArgs a{};
    a.base = v.base;
    a.N = N;
    a.dp = v.dp;
    a.nch = nch;
    a.knn = knn;
    a.kdeg = kdeg;
    a.kstride = kstride;
    a.R = R;
    a.C = pow2_ge(2 * R, 64);  // hit sort width
    a.lds_wave = (uint32_t)((((size_t)2 * R * 8 + 64 * 8 * 2 + 64 * 4 + (size_t)2 * R * 4 + (size_t)a.C * 4) + 15) &
                            ~(size_t)15);
    NSG_HIP(B.get(&a.pr_ids, (size_t)N * R));
    NSG_HIP(B.get(&a.pr_dist, (size_t)N * R));
    NSG_HIP(B.get(&a.pr_cnt, N));
    switch (cpl) {
    case 1: NSG_HIP(launch_dpg<1>(a, s)); break;
    case 2: NSG_HIP(launch_dpg<2>(a, s)); break;
    case 4: NSG_HIP(launch_dpg<4>(a, s)); break;
    case 8: NSG_HIP(launch_dpg<8>(a, s)); break;
    default: NSG_HIP(launch_dpg<16>(a, s)); break;
    }
    NSG_HIP(hipStreamSynchronize(s));
    const double prune_ms = ms_since(t_all);
    const auto t1 = std::chrono::high_resolution_clock::now();
    // union CSR: own count + in-degree per node
    uint32_t *incnt = nullptr, *tot = nullptr, *off = nullptr, *cur = nullptr, *ucnt = nullptr;
    NSG_HIP(B.get(&incnt, (size_t)N + 1));
    NSG_HIP(B.get(&tot, (size_t)N + 1));
    NSG_HIP(B.get(&off, (size_t)N + 1));
    NSG_HIP(B.get(&cur, N));
    NSG_HIP(B.get(&ucnt, N));
    NSG_HIP(hipMemsetAsync(incnt, 0, ((size_t)N + 1) * 4, s));
    NSG_HIP(hipMemsetAsync(tot, 0, ((size_t)N + 1) * 4, s));
    NSG_HIP(hipMemsetAsync(cur, 0, (size_t)N * 4, s));
    hipLaunchKernelGGL(nsg_rev_count_kernel, dim3(1024), dim3(256), 0, s, a.pr_ids, a.pr_cnt, N, R, incnt);
    NSG_HIP(hipGetLastError());
    std::vector<uint32_t> h_in(N + 1), h_own(N);
    NSG_HIP(hipMemcpyAsync(h_in.data(), incnt, (size_t)N * 4, hipMemcpyDeviceToHost, s));
    NSG_HIP(hipMemcpyAsync(h_own.data(), a.pr_cnt, (size_t)N * 4, hipMemcpyDeviceToHost, s));
    NSG_HIP(hipStreamSynchronize(s));
    std::vector<uint32_t> h_off(N + 1, 0);
    uint64_t tot64 = 0;
    for (uint32_t i = 0; i < N; ++i) {
        tot64 += (uint64_t)h_own[i] + h_in[i];
        if (tot64 >= (1ull << 31))
            return index_set_error(ix, WV_ENOTSUP, "DPG union above 2^31 edges (the edge offsets are 32-bit)");
        h_off[i + 1] = (uint32_t)tot64;
    }
    const uint32_t total = h_off[N];
    NSG_HIP(hipMemcpyAsync(off, h_off.data(), ((size_t)N + 1) * 4, hipMemcpyHostToDevice, s));
    uint64_t *keys = nullptr, *sorted = nullptr;
    NSG_HIP(B.get(&keys, total));
    NSG_HIP(B.get(&sorted, total));
    hipLaunchKernelGGL(dpg_union_fill_kernel, dim3(1024), dim3(256), 0, s, a.pr_ids, a.pr_dist, a.pr_cnt, N, R, off,
                       cur, keys);
    NSG_HIP(hipGetLastError());
    size_t bytes = 0;
    NSG_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, bytes, keys, sorted, (int)total, (int)N, off, off + 1,
                                                      0, 64, s));
    uint8_t *tmpb = nullptr;
    NSG_HIP(B.get(&tmpb, bytes));
    NSG_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(tmpb, bytes, keys, sorted, (int)total, (int)N, off, off + 1, 0,
                                                      64, s));
    hipLaunchKernelGGL(dpg_union_rows_kernel, dim3(8192), dim3(256), 0, s, sorted, off, N, ucnt, 0u, nullptr, nullptr,
                       nullptr, 0);
    NSG_HIP(hipGetLastError());
    NSG_HIP(hipStreamSynchronize(s));  // the context's stream does not wait for hipMemcpy
    std::vector<uint32_t> h_u(N);
    NSG_HIP(hipMemcpy(h_u.data(), ucnt, (size_t)N * 4, hipMemcpyDeviceToHost));
    uint32_t W = 1;
    for (uint32_t x : h_u) W = std::max(W, x);
    uint32_t *gid = nullptr, *gdeg = nullptr;
    float *gdist = nullptr;
    int rc = index_alloc_graph_dev(ix, W, &gid, &gdist, &gdeg);  // frees the kNN graph
    if (rc) return rc;
    hipLaunchKernelGGL(dpg_union_rows_kernel, dim3(8192), dim3(256), 0, s, sorted, off, N, ucnt, W, gid, gdist, gdeg,
                       1);
    NSG_HIP(hipGetLastError());
    NSG_HIP(hipStreamSynchronize(s));
    if (stats) {
        *stats = wv_nsg_stats{};
        stats->total_ms = ms_since(t_all);
        stats->prune_ms = prune_ms;
        stats->insert_ms = ms_since(t1);
        stats->width = W;
    }
    return WV_OK;
}

}  // namespace wv
