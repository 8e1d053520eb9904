// wv_search.hip -- greedy best-first search, one wavefront per query.
//
// Replaces ComponentSearchEntryRand::SearchEntryInner
// (src/component_search_entry.cpp:9-30) followed by
// ComponentSearchRouteGreedy::RouteInner (src/component_route.cpp:9-53),
// with Index::InsertIntoPool (include/weavess/index.h:152-180) done as a
// wave-wide merge.
//
// Per wavefront, in LDS: the sorted pool of L (dist,id) keys, its
// "unexpanded" flags, a sorted side buffer of at most B keys with flags, a
// 64-entry candidate block and an open-addressing visited table.  The
// search's candidate set is U = the L smallest keys of pool + buffer; the
// buffer only delays merges (with B = 0 candidates go straight into the
// pool).  One expansion:
//   1. ballots over the flags find the first unexpanded key of the pool and
//      of the buffer; the smaller one is expanded if fewer than L keys of
//      pool + buffer are smaller (it is in U), otherwise the search ends;
//   2. lane i takes neighbour i of that node (64 per block; more in blocks);
//   3. LDS atomicCAS test-and-set in the visited table; survivors are
//      compacted;
//   4. distances, 16 lanes per row with float4 gathers, in the build order;
//   5. candidates not below the L-th key of pool + buffer, or already in
//      either, are dropped; the rest are bitonic-sorted across the wave and
//      merged in place into the buffer (into the pool when B = 0); a full
//      buffer is first merged into the pool, keeping the L smallest.
// Selecting the L smallest keys is associative, so U, and with it the
// expansion order, is the same as inserting every candidate into an L-slot
// pool at once (Index::InsertIntoPool, include/weavess/index.h:152-180); the
// next expansion is always U's first unexpanded key, which is where the
// k-pointer restart of src/component_route.cpp:44-45 lands when (dist,id) is
// a strict order.
// The visited table forgets: when an id finds no free slot within the probe
// limit (8 slots until more ids than slots were inserted, 2 after) it evicts
// the occupant of its home slot.  A forgotten node that comes
// back is re-evaluated; it is then either already in pool or buffer
// (duplicate) or not below the L-th key, so ids and distances are unchanged
// and only the evaluation count grows.
#include "wv_device.hpp"
#include "wv_internal.hpp"
#include "wv_mlp.hpp"
#include "wv_pool.hpp"

#include <algorithm>

namespace wv {

namespace {

constexpr int VIS_PROBES = 8;      // probes while fewer ids than slots were inserted
constexpr int VIS_PROBES_FULL = 2; // after that

// Test-and-set of `id` in the visited table, probing at most `probes` slots.
// Returns 0 when the id was there, 1 when this call inserted it into a free
// slot, 2 when it found no free slot and evicted the occupant of its home
// slot (the table forgets).  Both 1 and 2 mean "evaluate".
__device__ inline int visit_first(uint32_t *vis, uint32_t hmask, uint32_t id, int probes) {
    const uint32_t h0 = fmix32(id) & hmask;
    uint32_t h = h0;
    for (int p = 0; p < probes; ++p) {
        const uint32_t old = atomicCAS(&vis[h], ID_NONE, id);
        if (old == ID_NONE) return 1;
        if (old == id) return 0;
        h = (h + 1) & hmask;
    }
    atomicExch(&vis[h0], id);
    return 2;
}

// First slot >= from of a[0..n) holding an unexpanded entry, or n; `key`
// gets that entry's key without the flag (KEY_EMPTY when there is none),
// taken from the lane that read it.
__device__ inline int first_flag(const uint64_t *a, int from, int n, uint64_t &key) {
    const int lane = lane_id();
    for (int t0 = from & ~(WAVE - 1); t0 < n; t0 += WAVE) {
        const int p = t0 + lane;
        const uint64_t e = p >= from && p < n ? a[p] : 0;
        const uint64_t b = __ballot((e & FBIT) != 0);
        if (b) {
            const int i = __ffsll((unsigned long long)b) - 1;
            key = canon(readlane64(e, i));
            return t0 + i;
        }
    }
    key = KEY_EMPTY;
    return n;
}

// first_flag over a[fa..na) and b[fb..nb) together: the first 64-slot chunk
// of each is read before either is tested, so the two LDS reads travel
// together; a search that does not end in its first chunk goes on as
// first_flag.  nb may be 0 (then ib = 0, kb = KEY_EMPTY).
__device__ inline void first_flag2(const uint64_t *a, int fa, int na, int &ia, uint64_t &ka, const uint64_t *b,
                                   int fb, int nb, int &ib, uint64_t &kb) {
    const int lane = lane_id();
    const int ta = fa & ~(WAVE - 1), tb = fb & ~(WAVE - 1);
    const int pa = ta + lane, pb = tb + lane;
    const uint64_t ea = pa >= fa && pa < na ? a[pa] : 0;
    const uint64_t eb = pb >= fb && pb < nb ? b[pb] : 0;
    const uint64_t ma = __ballot((ea & FBIT) != 0), mb = __ballot((eb & FBIT) != 0);
    if (ma) {
        const int i = __ffsll((unsigned long long)ma) - 1;
        ka = canon(readlane64(ea, i));
        ia = ta + i;
    } else {
        ia = first_flag(a, ta + WAVE, na, ka);
    }
    if (mb) {
        const int i = __ffsll((unsigned long long)mb) - 1;
        kb = canon(readlane64(eb, i));
        ib = tb + i;
    } else if (nb == 0) {
        kb = KEY_EMPTY;
        ib = 0;
    } else {
        ib = first_flag(b, tb + WAVE, nb, kb);
    }
}

// The k-th smallest (0-based) key of the union of the sorted, disjoint entry
// arrays a[0..na) and b[0..nb), k < na + nb: the split that takes i keys of b
// and k+1-i of a, with i the largest value for which b[i-1] < a[k+1-i].
// Every lane reads the same slots and returns the same (canonical) key.
__device__ inline uint64_t kth_union(const uint64_t *a, int na, const uint64_t *b, int nb, int k) {
    int lo = max(0, k + 1 - na), hi = min(nb, k + 1);
    while (lo < hi) {
        const int i = (lo + hi + 1) >> 1;
        const int j = k + 1 - i;
        if (j >= na || canon(b[i - 1]) < canon(a[j])) lo = i;
        else hi = i - 1;
    }
    const int j = k + 1 - lo;
    const uint64_t x = lo > 0 ? canon(b[lo - 1]) : 0, y = j > 0 ? canon(a[j - 1]) : 0;
    return x > y ? x : y;
}

// `key` if it enters U -- below the L-th key `worst` and in neither pool nor
// buffer (a duplicate comes from a forgotten node evaluated again) -- else
// KEY_EMPTY.  Both searches advance in one loop so their LDS reads overlap,
// 8-ary while a range holds 16 or more slots, then binary; worst <=
// pool[L-1], so the pool slot read is in range.
__device__ inline uint64_t survivor(uint64_t key, const uint64_t *pool, int Li, const uint64_t *buf, int nb,
                                    uint64_t worst) {
    if (!(key < worst)) return KEY_EMPTY;
    int plo = 0, phi = Li, blo = 0, bhi = nb;
    while (phi - plo >= 16 || bhi - blo >= 16) {
        if (phi - plo >= 16) narrow8_e(pool, plo, phi, key);
        if (bhi - blo >= 16) narrow8_e(buf, blo, bhi, key);
    }
    while (plo < phi || blo < bhi) {
        if (plo < phi) {
            const int mid = (plo + phi) >> 1;
            if (canon(pool[mid]) < key) plo = mid + 1;
            else phi = mid;
        }
        if (blo < bhi) {
            const int mid = (blo + bhi) >> 1;
            if (canon(buf[mid]) < key) blo = mid + 1;
            else bhi = mid;
        }
    }
    return (canon(pool[plo]) == key || (blo < nb && canon(buf[blo]) == key)) ? KEY_EMPTY : key;
}

// The survivors (lanes whose `key` is not KEY_EMPTY, mask lm, c of them,
// distinct keys) in ascending order into blk[0..c), and lane t < c's `key`
// becomes the t-th (lanes from c up get KEY_EMPTY).  Up to 16 survivors
// count their ranks from one another's registers and are written straight
// to their slots; more are compacted and sorted across the wavefront.
__device__ inline uint64_t sort_survivors(uint64_t key, uint64_t lm, int c, uint64_t *blk) {
    const int lane = lane_id();
    if (c <= 16) {
        int rank = 0;
        for (uint64_t mm = lm; mm; mm &= mm - 1) rank += readlane64(key, __ffsll((unsigned long long)mm) - 1) < key;
        wave_sync();
        if (key != KEY_EMPTY) blk[rank] = key;
        wave_sync();
        return lane < c ? blk[lane] : KEY_EMPTY;
    }
    wave_sync();
    if (key != KEY_EMPTY) blk[lanes_below(lm)] = key;
    wave_sync();
    int w = 2;
    while (w < c) w <<= 1;
    key = wave_sort64_n<true>(lane < c ? blk[lane] : KEY_EMPTY, w);
    wave_sync();
    if (lane < c) blk[lane] = key;
    wave_sync();
    return lane < c ? key : KEY_EMPTY;
}

// WATCH: the variant for the learned stop and per-hop traces (the state
// features after every expansion, the network in registers); plain launches
// use the variant without that code.
// One query, by one wavefront, in its LDS region `my` (every slot it reads
// is written first, so a region can serve query after query).
template <int CPL, int G, bool COLLECT, bool WATCH, bool LOOK = false>
__device__ inline void search_query(const SearchArgs &a, uint8_t *my, const uint32_t q) {
    const int lane = lane_id();
    uint64_t *pool = reinterpret_cast<uint64_t *>(my);         // P keys (P >= L, multiple of 64)
    uint64_t *ckey = pool + a.P;                                // 64
    uint32_t *cid = reinterpret_cast<uint32_t *>(ckey + WAVE);  // 64
    uint32_t *vis = cid + WAVE;                                 // H (a power of two >= 64)
    uint64_t *buf = reinterpret_cast<uint64_t *>(vis + a.H);    // B
    uint64_t *full = buf + a.B;                                 // C (collect mode)
    uint64_t *fblk = full + a.C;                                // 64 (collect mode)
    uint64_t *fstg = fblk + WAVE;                               // 128 (collect mode)
    int flen = 0, fsn = 0;

    const uint32_t N = a.N, nch = a.nch, dp = a.dp;
    // per-query budgets were checked on the host when the launch was planned;
    // clamping to the plan keeps a buffer rewritten since inside the pool
    const uint32_t lcap = min(a.P, a.entries ? N : (N > 1 ? N - 1 : 1u));
    const uint32_t L = min(max(a.L_per_q ? a.L_per_q[q] : a.L, 1u), lcap);
    const uint32_t kq = min(max(a.k_per_q ? a.k_per_q[q] : a.k, 1u), min(L, a.k_stride));

    float4 qv[CPL];
    load_query<CPL>(a.queries + (size_t)q * dp, nch, qv);

    // ---- entries (SearchEntryRand): not marked visited, all flagged.  The
    // pool's storage first holds the L entry ids (one per 8-byte slot), then
    // their keys: distances 64 at a time, each sorted chunk merged into the
    // growing pool.  A merge writes only slots below the merged length, and
    // the ids of later chunks sit above it.
    if (a.entries) {
        for (uint32_t i = lane; i < L; i += WAVE) {
            const uint32_t e = a.entries[(size_t)q * L + i];
            pool[i] = e < N ? e : N - 1;  // out-of-range ids (unchecked device input) are clamped
        }
        wave_sync();
    } else {
        gen_random_lds(stream_of(a.seed, TAG_ENTRY, q), L, N, pool, ckey);
    }
    {
        int len = 0;
        for (uint32_t c0 = 0; c0 < L; c0 += WAVE) {
            const int cnt = (int)min((uint32_t)WAVE, L - c0);
            if (lane < cnt) cid[lane] = (uint32_t)pool[c0 + lane];
            wave_sync();
            rows_dist<CPL, G, true>(a.base, dp, nch, qv, cid, cnt, ckey);
            wave_sync();
            if (COLLECT) collect_stage(full, flen, (int)a.C, ckey, cnt, fstg, fsn, fblk);
            const uint64_t key = wave_sort64<true>(lane < cnt ? ckey[lane] : KEY_EMPTY);
            wave_sync();
            ckey[lane] = key;
            wave_sync();
            merge_into(pool, len, (int)L, ckey, cnt, true);  // entries start unexpanded
            len += cnt;
        }
    }
    for (uint32_t i = L + lane; i < a.P; i += WAVE) pool[i] = KEY_EMPTY;
    const uint32_t hmask = a.H - 1;
    for (uint32_t i = lane; i < a.H; i += WAVE) vis[i] = ID_NONE;
    wave_sync();

    uint32_t evals = L, hops = 0;
    const int Li = (int)L;
    const int B = (int)a.B;    // buffer capacity; 0: merge straight into the pool
    int kp = 0;                // pool slots before kp are expanded
    int nb = 0, kb = 0;        // buffer length; buffer slots before kb are expanded
    uint64_t worst = canon(pool[Li - 1]);  // the L-th key of pool + buffer
    // learned termination / traces: state features (identical on every lane)
    constexpr bool watch = WATCH;
    TermRegs net;
    const bool use_model = WATCH && a.model != nullptr;
    if (use_model) term_load(a.model, net);
    // the settle hop is also an output of its own (without traces or a model)
    const bool track_settle = watch || a.settle != nullptr;
    const int kidx = (int)(kq < L ? kq : L) - 1;
    const float entry_d = watch ? key_dist(pool[0]) : 0.f;
    // k-th distance 1..4 expansions ago
    float r0 = watch ? key_dist(pool[kidx]) : 0.f, r1 = r0, r2 = r0, r3 = r0;  // bit 31 is not in the distance
    float kprev = r0;  // k-th distance at the last stop-rule evaluation (the entries' at first)
    uint32_t settle = 0;  // last expansion that changed the top-k
    int probes = VIS_PROBES;
    if constexpr (LOOK) {
        // The plain loop for graphs of at most 64 neighbour slots a node (one
        // block a node), one step further ahead: once the next node is
        // chosen, its neighbours are tested in the visited table and their
        // rows' loads issued before this node's survivors are merged, so the
        // gathers travel during the merge.  The visited table and the merge
        // touch disjoint state and each keeps its order of operations, so
        // expansions, evaluations and results are the loop's below.
        static_assert(!COLLECT && !WATCH, "the look-ahead loop is the plain search's");
        uint32_t id0 = ID_NONE, du = 0;
        auto start = [&](uint32_t node) {
            id0 = (uint32_t)lane < a.deg_max ? a.graph[(size_t)node * a.deg_max + lane] : ID_NONE;
            du = a.deg[node];
        };
        int m = 0;
        float4 v[G][CPL];
        auto mark = [&]() {  // step 1 of the chosen node, then its first rows' loads
            const uint32_t d = du < a.deg_max ? du : a.deg_max;
            const uint32_t id = (uint32_t)lane < d ? id0 : ID_NONE;
            const bool fresh = id < N && visit_first(vis, hmask, id, probes) != 0;
            const uint64_t bm = __ballot(fresh);
            m = __popcll(bm);
            if (fresh) cid[lanes_below(bm)] = id;
            wave_sync();
            if (m > 0) rows_issue<CPL, G>(a.base, dp, nch, cid, m, 0, v);
        };
        bool live;
        {
            uint64_t cur;
            const int kk = first_flag(pool, kp, Li, cur);  // the buffer is empty
            live = !(cur == KEY_EMPTY || cur > worst);
            if (live) {
                wave_sync();
                if (lane == 0) pool[kk] = cur;  // expanded: stored without bit 31
                kp = kk + 1;
                start(key_id(cur));
                mark();
            }
        }
        while (live) {
            ++hops;
            int c = 0;
            uint64_t key = KEY_EMPTY;
            if (m > 0) {
                rows_finish<CPL, G, true>(a.base, dp, nch, qv, cid, m, v, ckey);
                evals += (uint32_t)m;
                if (evals - L > a.H) probes = VIS_PROBES_FULL;
                wave_sync();
                key = survivor(lane < m ? ckey[lane] : KEY_EMPTY, pool, Li, buf, nb, worst);
                const uint64_t lm = __ballot(key != KEY_EMPTY);
                c = __popcll(lm);
                if (c > 0) key = sort_survivors(key, lm, c, ckey);
            }
            bool next_new;
            {
                uint64_t ka, kbk;
                int kk, jb;
                first_flag2(pool, kp, Li, kk, ka, buf, kb, nb, jb, kbk);
                const bool from_buf = kbk < ka;
                const uint64_t x = from_buf ? kbk : ka;
                const uint64_t y = c > 0 ? readlane64(key, 0) : KEY_EMPTY;
                next_new = y < x;
                const uint64_t cur = next_new ? y : x;
                live = !(cur == KEY_EMPTY || cur > worst);
                if (live) {
                    if (!next_new) {
                        wave_sync();
                        if (lane == 0) {
                            if (from_buf) buf[jb] = cur;
                            else pool[kk] = cur;
                        }
                        if (from_buf) kb = jb + 1;
                        else kp = kk + 1;
                    }
                    start(key_id(cur));
                    mark();
                }
            }
            if (c == 0) continue;
            if (B == 0) {
                const int pos0 = merge_into(pool, Li, Li, ckey, c, true);
                if (pos0 < kp) kp = pos0;
                if (pos0 <= kidx) settle = hops;
                if (next_new && lane == 0) pool[pos0] = ckey[0];  // pos0 < L: ckey[0] < worst
                worst = canon(pool[Li - 1]);
            } else {
                if (nb + c > B) {
                    const int pos0 = merge_into(pool, Li, Li, buf, nb, false);
                    if (pos0 < kp) kp = pos0;
                    nb = 0;
                    kb = 0;
                }
                const int bp = buffer_insert(buf, nb, key, c, pool, Li, worst);
                nb += c;
                if (bp < kb) kb = bp;
                if (next_new && lane == 0) buf[bp] = ckey[0];
                if (track_settle && bp + lower_bound_e(pool, Li, ckey[0]) <= kidx) settle = hops;
            }
            wave_sync();
        }
    } else if constexpr (!COLLECT) {
        // The loop below with the next node chosen before the merge: once
        // the last block of u's survivors is sorted, the next node is the
        // smaller of U's first unexpanded key and the smallest survivor (both
        // are in U after the merge, which keeps every other key's order).
        // It is marked expanded and its degree and first neighbour slots are
        // loaded before the block is merged, so those loads travel during
        // the merge.  Expansions, evaluations and results are the loop's.
        uint32_t u = 0, du = 0, id0 = ID_NONE;
        auto start = [&](uint32_t node) {
            // a row holds deg_max slots, so the first 64 load with the
            // degree (those past it are in bounds and masked off)
            u = node;
            const uint32_t *adj = a.graph + (size_t)node * a.deg_max;
            id0 = (uint32_t)lane < a.deg_max ? adj[lane] : ID_NONE;
            du = a.deg[node];
        };
        bool live;
        {
            uint64_t cur;
            const int kk = first_flag(pool, kp, Li, cur);  // the buffer is empty
            live = !(cur == KEY_EMPTY || cur > worst);
            if (live) {
                wave_sync();
                if (lane == 0) pool[kk] = cur;  // expanded: stored without bit 31
                kp = kk + 1;
                start(key_id(cur));
            }
        }
        while (live) {
            ++hops;
            if (du > a.deg_max) du = a.deg_max;
            const uint32_t nblk = du > (uint32_t)WAVE ? (du + WAVE - 1) / WAVE : 1u;
            const uint32_t *adj = a.graph + (size_t)u * a.deg_max;
            const uint32_t first = id0;
            for (uint32_t bi = 0; bi < nblk; ++bi) {
                const uint32_t mi = bi * WAVE + lane;
                const uint32_t id = mi < du ? (bi == 0 ? first : adj[mi]) : ID_NONE;
                const bool fresh = id < N && visit_first(vis, hmask, id, probes) != 0;
                const uint64_t bm = __ballot(fresh);
                const int m = __popcll(bm);
                if (fresh) cid[lanes_below(bm)] = id;
                wave_sync();
                int c = 0;
                uint64_t key = KEY_EMPTY;
                if (m > 0) {
                    rows_dist<CPL, G, true>(a.base, dp, nch, qv, cid, m, ckey);
                    evals += (uint32_t)m;
                    if (evals - L > a.H) probes = VIS_PROBES_FULL;
                    wave_sync();
                    key = survivor(lane < m ? ckey[lane] : KEY_EMPTY, pool, Li, buf, nb, worst);
                    const uint64_t lm = __ballot(key != KEY_EMPTY);
                    c = __popcll(lm);
                    if (c > 0) key = sort_survivors(key, lm, c, ckey);
                }
                bool next_new = false;  // the next node is the smallest survivor
                if (bi + 1 == nblk) {
                    uint64_t ka, kbk;
                    int kk, jb;
                    first_flag2(pool, kp, Li, kk, ka, buf, kb, nb, jb, kbk);
                    const bool from_buf = kbk < ka;
                    const uint64_t x = from_buf ? kbk : ka;
                    const uint64_t y = c > 0 ? readlane64(key, 0) : KEY_EMPTY;
                    next_new = y < x;
                    const uint64_t cur = next_new ? y : x;
                    // with survivors cur <= y < worst; without, worst stays
                    live = !(cur == KEY_EMPTY || cur > worst);
                    if (live) {
                        if (!next_new) {
                            wave_sync();
                            if (lane == 0) {
                                if (from_buf) buf[jb] = cur;
                                else pool[kk] = cur;
                            }
                            if (from_buf) kb = jb + 1;
                            else kp = kk + 1;
                        }
                        start(key_id(cur));
                    }
                }
                if (c == 0) continue;
                if (B == 0) {
                    const int pos0 = merge_into(pool, Li, Li, ckey, c, true);
                    if (pos0 < kp) kp = pos0;
                    if (pos0 <= kidx) settle = hops;
                    if (next_new && lane == 0) pool[pos0] = ckey[0];  // pos0 < L: ckey[0] < worst
                    worst = canon(pool[Li - 1]);
                } else {
                    if (nb + c > B) {
                        const int pos0 = merge_into(pool, Li, Li, buf, nb, false);
                        if (pos0 < kp) kp = pos0;
                        nb = 0;
                        kb = 0;
                    }
                    const int bp = buffer_insert(buf, nb, key, c, pool, Li, worst);
                    nb += c;
                    if (bp < kb) kb = bp;
                    if (next_new && lane == 0) buf[bp] = ckey[0];
                    if (track_settle && bp + lower_bound_e(pool, Li, ckey[0]) <= kidx) settle = hops;
                }
                wave_sync();
            }
            if constexpr (WATCH) {
                // The state after `hops` expansions: features and the trace
                // row after every expansion while tracing, the stop rule after
                // every MLP_STRIDE-th.  The node already chosen for the next
                // hop is expanded only if the rule does not stop here
                // (choosing it changed no key's order, only its flag).
                const bool tr = a.trace != nullptr && hops <= a.trace_cap;
                const bool at = hops % MLP_STRIDE == 0;
                if (tr || (use_model && at)) {
                    float f[MLP_IN];
                    const uint64_t kth = nb ? kth_union(pool, Li, buf, nb, kidx) : canon(pool[kidx]);
                    const uint64_t best = (nb && canon(buf[0]) < canon(pool[0])) ? canon(buf[0]) : canon(pool[0]);
                    // k-th distance MLP_STRIDE expansions ago: the per-hop
                    // history while tracing, else the last evaluation's
                    term_features((float)hops, key_dist(kth), key_dist(best), entry_d, tr ? r3 : kprev,
                                  (float)(hops - settle), (float)L, (float)kq, f);
                    if (tr) {
                        r3 = r2;
                        r2 = r1;
                        r1 = r0;
                        r0 = key_dist(kth);
                        if (lane == 0) {
                            float4 *t = reinterpret_cast<float4 *>(a.trace +
                                                                   ((size_t)q * a.trace_cap + (hops - 1)) * MLP_IN);
                            t[0] = make_float4(f[0], f[1], f[2], f[3]);
                            t[1] = make_float4(f[4], f[5], f[6], f[7]);
                        }
                    }
                    if (at) kprev = key_dist(kth);
                    if (at && live && use_model && hops >= net.min_hops && term_logit_regs(net, f) > net.threshold)
                        break;
                }
            }
        }
    } else {
        static_assert(!WATCH, "the collecting mode runs without the stop rule or traces");
        for (;;) {
            // U's first unexpanded key: the smaller of the pool's and the buffer's
            uint64_t ka, kbk = KEY_EMPTY;
            const int kk = first_flag(pool, kp, Li, ka);
            const int jb = nb ? first_flag(buf, kb, nb, kbk) : 0;
            const bool from_buf = kbk < ka;
            const uint64_t cur = from_buf ? kbk : ka;
            // cur is in U iff it is not above U's L-th key (`worst` is kept exact)
            const bool done = cur == KEY_EMPTY || cur > worst;
            if (done) break;
            const uint32_t u = key_id(cur);  // cur is canonical
            wave_sync();
            if (lane == 0) {  // expanded: store the key without bit 31
                if (from_buf) buf[jb] = cur;
                else pool[kk] = cur;
            }
            if (from_buf) kb = jb + 1;
            else kp = kk + 1;
            ++hops;
            // the first 64 neighbour slots load together with the degree, not
            // after it: a row holds deg_max slots, so those past the degree are
            // in bounds, and they are masked off below
            const uint32_t *adj = a.graph + (size_t)u * a.deg_max;
            const uint32_t id0 = (uint32_t)lane < a.deg_max ? adj[lane] : ID_NONE;
            uint32_t du = a.deg[u];
            if (du > a.deg_max) du = a.deg_max;
            for (uint32_t m0 = 0; m0 < du; m0 += WAVE) {
                const uint32_t mi = m0 + lane;
                const uint32_t id = mi < du ? (m0 == 0 ? id0 : adj[mi]) : ID_NONE;
                const bool fresh = id < N && visit_first(vis, hmask, id, probes) != 0;
                const uint64_t bm = __ballot(fresh);
                const int m = __popcll(bm);
                if (fresh) cid[lanes_below(bm)] = id;
                wave_sync();
                if (m == 0) continue;
                rows_dist<CPL, G, true>(a.base, dp, nch, qv, cid, m, ckey);
                evals += (uint32_t)m;
                // more ids inserted than the table has slots: it forgets either
                // way, so stop probing deeply
                if (evals - L > a.H) probes = VIS_PROBES_FULL;
                wave_sync();
                if (COLLECT) collect_stage(full, flen, (int)a.C, ckey, m, fstg, fsn, fblk);
                // the batch's keys are distinct (the visited table admits an id
                // once per batch)
                uint64_t key = survivor(lane < m ? ckey[lane] : KEY_EMPTY, pool, Li, buf, nb, worst);
                const uint64_t lm = __ballot(key != KEY_EMPTY);
                if (lm == 0) continue;
                const int c = __popcll(lm);
                key = sort_survivors(key, lm, c, ckey);
                if (B == 0) {
                    const int pos0 = merge_into(pool, Li, Li, ckey, c, true);
                    if (pos0 < kp) kp = pos0;
                    if (pos0 <= kidx) settle = hops;  // a candidate entered the top-k
                    worst = canon(pool[Li - 1]);
                } else {
                    if (nb + c > B) {  // full: merge the buffer into the pool
                        const int pos0 = merge_into(pool, Li, Li, buf, nb, false);
                        if (pos0 < kp) kp = pos0;
                        nb = 0;
                        kb = 0;
                    }
                    const int bp = buffer_insert(buf, nb, key, c, pool, Li, worst);
                    nb += c;
                    if (bp < kb) kb = bp;
                    if (track_settle && bp + lower_bound_e(pool, Li, ckey[0]) <= kidx) settle = hops;
                }
            }
        }
    }
    if (nb) merge_into(pool, Li, Li, buf, nb, false);
    if (COLLECT) {
        flen = collect_flush(full, flen, (int)a.C, fstg, fsn, fblk);
        wave_sync();
        for (uint32_t i = lane; i < a.C; i += WAVE) a.full[(size_t)q * a.C + i] = (int)i < flen ? full[i] : KEY_EMPTY;
    }

    for (uint32_t i = lane; i < a.k_stride; i += WAVE) {
        const bool ok = i < kq;
        a.out_ids[(size_t)q * a.k_stride + i] = ok ? entry_id(pool[i]) : ID_NONE;
        if (a.out_dist) a.out_dist[(size_t)q * a.k_stride + i] = ok ? key_dist(pool[i]) : -1.0f;
    }
    if (lane == 0) {
        if (a.evals) a.evals[q] = evals;
        if (a.hops) a.hops[q] = hops;
        if (a.trace_len) a.trace_len[q] = hops < a.trace_cap ? hops : a.trace_cap;
        if (a.settle) a.settle[q] = settle;
    }
}

// Wavefronts per SIMD the plain grid kernel is compiled for: 3 (at most 168
// VGPRs, one or two values spilled) for the 64-row gather tile of rows of
// 17-24 chunks, which otherwise takes a few more and drops to 2 where LDS
// would allow 3.  The look-ahead loop is compiled for 2.
// (The persistent form keeps more state live and would spill under it.)
template <int CPL, int G, bool EXTRA> struct SearchWaves {
    static constexpr int value = EXTRA ? 1
                               : (CPL == 3 && G == 8) ? 3
                               : ((CPL == 1 && G == 8) || (CPL == 2 && G == 4) || (CPL == 3 && G == 2)) ? 5 : 1;
};

// One wavefront per query: query = its global wavefront index.
template <int CPL, int G, bool COLLECT, bool WATCH, bool LOOK = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LOOK ? 2 : SearchWaves<CPL, G, COLLECT || WATCH>::value)))
greedy_search_kernel(SearchArgs a) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int wid = (int)(threadIdx.x >> 6);
    const uint32_t q = blockIdx.x * (blockDim.x >> 6) + wid;
    if (q >= a.Q) return;
    search_query<CPL, G, COLLECT, WATCH, LOOK>(a, smem + (size_t)wid * a.wave_lds, q);
}

// Persistent form: a grid of about as many wavefronts as the GPU holds at
// once, each taking query slots from the device work queue `a.queue` (a
// counter the host zeroes before the launch) until the slots run out.  Lane
// 0 claims a slot with one atomicAdd and broadcasts it; a slot at or past Q
// ends the wavefront, and the counter only grows, so every wavefront ends
// once the Q slots are claimed.  Slot i serves query order[i] (longest
// budgets first) or query i.
template <int CPL, int G, bool WATCH>
__global__ void __launch_bounds__(256) greedy_search_persistent_kernel(SearchArgs a) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int wid = (int)(threadIdx.x >> 6);
    const int lane = lane_id();
    uint8_t *my = smem + (size_t)wid * a.wave_lds;
    for (;;) {
        uint32_t slot = 0;
        if (lane == 0) slot = atomicAdd(a.queue, 1u);
        slot = (uint32_t)__shfl((int)slot, 0);
        if (slot >= a.Q) break;
        const uint32_t q = a.order ? a.order[slot] : slot;
        search_query<CPL, G, false, WATCH>(a, my, q);
        wave_sync();  // the next query rewrites this region
    }
}

template <int CPL, int G, bool COLLECT, bool WATCH, bool LOOK = false>
hipError_t launch_w(const SearchArgs &a, int waves_per_block, hipStream_t s) {
    const size_t lds = (size_t)a.wave_lds * waves_per_block;
    const uint32_t blocks = (a.Q + waves_per_block - 1) / waves_per_block;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(
            reinterpret_cast<const void *>(&greedy_search_kernel<CPL, G, COLLECT, WATCH, LOOK>),
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((greedy_search_kernel<CPL, G, COLLECT, WATCH, LOOK>), dim3(blocks),
                       dim3(WAVE * waves_per_block), lds, s, a);
    return hipGetLastError();
}
// ---------------------------------------------------------------- cooperative
// W wavefronts of one workgroup serve one query.  Wave 0 owns the pool, the
// merge buffer and the expansion order; every wave takes a share of the
// expanded node's neighbours (wave w the w-th 64 of each W x 64), tests them
// in the shared visited table and computes their distances into its own
// candidate block; wave 0 then filters, sorts and merges block after block,
// in neighbour order, exactly as one wavefront merges its 64-neighbour
// chunks, so the pool and the expansion order are the same.  The entries'
// distances are split the same way.  Workgroup barriers separate the
// phases; every wave passes the same barriers because the loop bounds and
// the stop decision are read from LDS after a barrier.
template <int CPL, int G, int W>
__global__ void __launch_bounds__(256) greedy_search_coop_kernel(SearchArgs a) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int wid = (int)(threadIdx.x >> 6);
    const int lane = lane_id();
    const uint32_t q = blockIdx.x;
    if (q >= a.Q) return;  // the whole workgroup
    uint64_t *pool = reinterpret_cast<uint64_t *>(smem);          // P
    uint64_t *buf = pool + a.P;                                    // B
    uint64_t *ckeys = buf + a.B;                                   // W x 64
    uint32_t *cids = reinterpret_cast<uint32_t *>(ckeys + W * WAVE);  // W x 64
    uint32_t *vis = cids + W * WAVE;                               // H
    // [0], [1] the node to expand or ID_NONE (alternating by hop, so wave 0
    // never overwrites a value another wave has yet to read), [2] evals,
    // [3..3+W) the waves' candidate counts
    uint32_t *ctl = vis + a.H;
    uint64_t *ckey = ckeys + wid * WAVE;
    uint32_t *cid = cids + wid * WAVE;

    const uint32_t N = a.N, nch = a.nch, dp = a.dp;
    const uint32_t lcap = min(a.P, a.entries ? N : (N > 1 ? N - 1 : 1u));
    const uint32_t L = min(max(a.L_per_q ? a.L_per_q[q] : a.L, 1u), lcap);
    const uint32_t kq = min(max(a.k_per_q ? a.k_per_q[q] : a.k, 1u), min(L, a.k_stride));
    float4 qv[CPL];
    load_query<CPL>(a.queries + (size_t)q * dp, nch, qv);
    // ---- entries: ids into the pool's storage (wave 0), then each round
    // every wave scores one 64-id chunk and sorts it in its block, and wave 0
    // merges the round's chunks in order (a merge writes only below the
    // merged length; later rounds' ids sit above it)
    if (wid == 0) {
        if (a.entries) {
            for (uint32_t i = lane; i < L; i += WAVE) {
                const uint32_t e = a.entries[(size_t)q * L + i];
                pool[i] = e < N ? e : N - 1;
            }
            wave_sync();
        } else {
            gen_random_lds(stream_of(a.seed, TAG_ENTRY, q), L, N, pool, ckey);
        }
    }
    __syncthreads();
    int len = 0;
    for (uint32_t r0 = 0; r0 < L; r0 += W * WAVE) {
        const uint32_t c0 = r0 + (uint32_t)wid * WAVE;
        if (c0 < L) {
            const int cnt = (int)min((uint32_t)WAVE, L - c0);
            if (lane < cnt) cid[lane] = (uint32_t)pool[c0 + lane];
            wave_sync();
            rows_dist<CPL, G, true>(a.base, dp, nch, qv, cid, cnt, ckey);
            wave_sync();
            const uint64_t key = wave_sort64<true>(lane < cnt ? ckey[lane] : KEY_EMPTY);
            wave_sync();
            ckey[lane] = key;
            wave_sync();
        }
        __syncthreads();
        if (wid == 0) {
            for (int w = 0; w < W; ++w) {
                const uint32_t cw = r0 + (uint32_t)w * WAVE;
                if (cw >= L) break;
                const int cnt = (int)min((uint32_t)WAVE, L - cw);
                merge_into(pool, len, (int)L, ckeys + w * WAVE, cnt, true);
                len += cnt;
            }
        }
        __syncthreads();
    }
    if (wid == 0)
        for (uint32_t i = L + lane; i < a.P; i += WAVE) pool[i] = KEY_EMPTY;
    const uint32_t hmask = a.H - 1;
    for (uint32_t i = threadIdx.x; i < a.H; i += blockDim.x) vis[i] = ID_NONE;
    if (threadIdx.x == 0) ctl[2] = L;
    __syncthreads();

    uint32_t evals = L, hops = 0;  // wave 0's
    uint32_t step = 0;             // every wave's count of selection steps
    const int Li = (int)L;
    const int B = (int)a.B;
    int kp = 0, nb = 0, kb = 0;
    uint64_t worst = canon(pool[Li - 1]);
    for (;;) {
        if (wid == 0) {
            uint64_t ka, kbk = KEY_EMPTY;
            const int kk = first_flag(pool, kp, Li, ka);
            const int jb = nb ? first_flag(buf, kb, nb, kbk) : 0;
            const bool from_buf = kbk < ka;
            const uint64_t cur = from_buf ? kbk : ka;
            const bool done = cur == KEY_EMPTY || cur > worst;
            wave_sync();
            if (!done) {
                if (lane == 0) {
                    if (from_buf) buf[jb] = cur;
                    else pool[kk] = cur;
                }
                if (from_buf) kb = jb + 1;
                else kp = kk + 1;
                ++hops;
            }
            if (lane == 0) ctl[step & 1] = done ? ID_NONE : key_id(cur);
        }
        __syncthreads();
        const uint32_t u = ctl[step & 1];
        ++step;
        if (u == ID_NONE) break;
        // probes as one wavefront would choose them from the evaluations so far
        const int probes = ctl[2] - L > a.H ? VIS_PROBES_FULL : VIS_PROBES;
        const uint32_t *adj = a.graph + (size_t)u * a.deg_max;
        uint32_t du = a.deg[u];
        if (du > a.deg_max) du = a.deg_max;
        for (uint32_t m0 = 0; m0 < du; m0 += W * WAVE) {
            const uint32_t mi = m0 + (uint32_t)wid * WAVE + lane;
            const uint32_t id = mi < du ? adj[mi] : ID_NONE;
            const bool fresh = id < N && visit_first(vis, hmask, id, probes) != 0;
            const uint64_t bm = __ballot(fresh);
            const int m = __popcll(bm);
            if (fresh) cid[lanes_below(bm)] = id;
            wave_sync();
            if (m > 0) rows_dist<CPL, G, true>(a.base, dp, nch, qv, cid, m, ckey);
            if (lane == 0) ctl[3 + wid] = (uint32_t)m;
            __syncthreads();
            if (wid == 0) {
                for (int w = 0; w < W; ++w) {
                    const int mw = (int)ctl[3 + w];
                    if (mw == 0) continue;
                    uint64_t *blk = ckeys + w * WAVE;
                    evals += (uint32_t)mw;
                    uint64_t key = lane < mw ? blk[lane] : KEY_EMPTY;
                    if (key < worst) {
                        int plo = 0, phi = Li, blo = 0, bhi = nb;
                        while (plo < phi || blo < bhi) {
                            if (plo < phi) {
                                const int mid = (plo + phi) >> 1;
                                if (canon(pool[mid]) < key) plo = mid + 1;
                                else phi = mid;
                            }
                            if (blo < bhi) {
                                const int mid = (blo + bhi) >> 1;
                                if (canon(buf[mid]) < key) blo = mid + 1;
                                else bhi = mid;
                            }
                        }
                        if (canon(pool[plo]) == key || (blo < nb && canon(buf[blo]) == key)) key = KEY_EMPTY;
                    } else {
                        key = KEY_EMPTY;
                    }
                    const uint64_t lm = __ballot(key != KEY_EMPTY);
                    if (lm == 0) continue;
                    const int c = __popcll(lm);
                    key = sort_survivors(key, lm, c, blk);
                    if (B == 0) {
                        const int pos0 = merge_into(pool, Li, Li, blk, c, true);
                        if (pos0 < kp) kp = pos0;
                        worst = canon(pool[Li - 1]);
                    } else {
                        if (nb + c > B) {
                            const int pos0 = merge_into(pool, Li, Li, buf, nb, false);
                            if (pos0 < kp) kp = pos0;
                            nb = 0;
                            kb = 0;
                        }
                        const int bp = buffer_insert(buf, nb, key, c, pool, Li, worst);
                        nb += c;
                        if (bp < kb) kb = bp;
                    }
                }
                if (lane == 0) ctl[2] = evals;
            }
            __syncthreads();  // the blocks and counts are rewritten next
        }
    }
    if (wid != 0) return;
    if (nb) merge_into(pool, Li, Li, buf, nb, false);
    for (uint32_t i = lane; i < a.k_stride; i += WAVE) {
        const bool ok = i < kq;
        a.out_ids[(size_t)q * a.k_stride + i] = ok ? entry_id(pool[i]) : ID_NONE;
        if (a.out_dist) a.out_dist[(size_t)q * a.k_stride + i] = ok ? key_dist(pool[i]) : -1.0f;
    }
    if (lane == 0) {
        if (a.evals) a.evals[q] = evals;
        if (a.hops) a.hops[q] = hops;
        if (a.settle) a.settle[q] = 0;
    }
}

template <int CPL, int G>
hipError_t launch_c(const SearchArgs &a, int wpq, hipStream_t s) {
    const size_t lds = search_coop_lds(a.P, a.H, a.B, (uint32_t)wpq);
    const void *fn = wpq == 4 ? reinterpret_cast<const void *>(&greedy_search_coop_kernel<CPL, G, 4>)
                              : reinterpret_cast<const void *>(&greedy_search_coop_kernel<CPL, G, 2>);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    if (wpq == 4)
        hipLaunchKernelGGL((greedy_search_coop_kernel<CPL, G, 4>), dim3(a.Q), dim3(4 * WAVE), lds, s, a);
    else
        hipLaunchKernelGGL((greedy_search_coop_kernel<CPL, G, 2>), dim3(a.Q), dim3(2 * WAVE), lds, s, a);
    return hipGetLastError();
}

__global__ void queue_reset_kernel(uint32_t *queue) {
    if (threadIdx.x == 0) *queue = 0u;
}

template <int CPL, int G, bool WATCH>
hipError_t launch_p(const SearchArgs &a, int waves_per_block, uint32_t resident_blocks, hipStream_t s) {
    const size_t lds = (size_t)a.wave_lds * waves_per_block;
    const void *fn = reinterpret_cast<const void *>(&greedy_search_persistent_kernel<CPL, G, WATCH>);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    // as many workgroups as are resident at once (fewer when Q is small)
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, WAVE * waves_per_block, lds);
    if (e != hipSuccess) return e;
    const uint32_t want = (a.Q + waves_per_block - 1) / waves_per_block;
    uint32_t blocks = (uint32_t)std::max(per_cu, 1) * resident_blocks;
    if (blocks > want) blocks = want;
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(64), 0, s, a.queue);
    hipLaunchKernelGGL((greedy_search_persistent_kernel<CPL, G, WATCH>), dim3(blocks), dim3(WAVE * waves_per_block),
                       lds, s, a);
    return hipGetLastError();
}

template <int CPL, int G, bool COLLECT>
hipError_t launch_t(const SearchArgs &a, int waves_per_block, uint32_t cu_count, hipStream_t s) {
    const bool watch = a.model || a.trace;
    if (!COLLECT && a.queue) {
        return watch ? launch_p<CPL, G, true>(a, waves_per_block, cu_count, s)
                     : launch_p<CPL, G, false>(a, waves_per_block, cu_count, s);
    }
    if (!COLLECT && watch) return launch_w<CPL, G, false, true>(a, waves_per_block, s);
    if constexpr (!COLLECT && !(CPL == 4 && G == 8)) {
        // The look-ahead loop holds the next node's rows in registers through
        // the merge (123-248 VGPRs by tile, 200 for rows of 17-24 chunks in
        // 64-row tiles; the 4-chunk 8-group tile would spill and keeps the
        // plain loop): taken for graphs of at most 64 slots a node when the
        // LDS already keeps the launch at 2 wavefronts a SIMD or fewer.  At
        // 2M x 96, L = 2075 it is 2.6 % faster than the loop without it.
        const size_t wg_lds = (size_t)a.wave_lds * waves_per_block;
        const size_t lds_waves = wg_lds ? (LDS_PER_CU / wg_lds) * waves_per_block : 0;
        if (a.deg_max <= (uint32_t)WAVE && lds_waves <= 8)
            return launch_w<CPL, G, false, false, true>(a, waves_per_block, s);
    }
    return launch_w<CPL, G, COLLECT, false>(a, waves_per_block, s);
}

}  // namespace

size_t search_wave_lds(uint32_t P, uint32_t H, uint32_t B, uint32_t C) {
    size_t b = (size_t)P * 8 + WAVE * 8 + WAVE * 4 + (size_t)H * 4 + (size_t)B * 8;
    if (C) b += (size_t)C * 8 + WAVE * 8 + 2 * WAVE * 8;
    return (b + 15) & ~(size_t)15;
}

size_t search_coop_lds(uint32_t P, uint32_t H, uint32_t B, uint32_t W) {
    const size_t b = (size_t)P * 8 + (size_t)B * 8 + (size_t)W * WAVE * 12 + (size_t)H * 4 + 64;
    return (b + 15) & ~(size_t)15;
}

hipError_t launch_greedy_search_coop(const SearchArgs &a, int chunks_per_lane, int waves_per_query, hipStream_t s) {
    if (a.full || a.model || a.trace || a.queue || (waves_per_query != 2 && waves_per_query != 4))
        return hipErrorInvalidValue;
    switch (chunks_per_lane) {
    case 1: return launch_c<1, 8>(a, waves_per_query, s);
    case 2: return launch_c<2, 4>(a, waves_per_query, s);
    case 3: return launch_c<3, 2>(a, waves_per_query, s);
    case 4: return launch_c<4, 4>(a, waves_per_query, s);
    case 8: return launch_c<8, 2>(a, waves_per_query, s);
    case 16: return launch_c<16, 1>(a, waves_per_query, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_greedy_search(const SearchArgs &a, int chunks_per_lane, int row_groups,
                                int waves_per_block, hipStream_t s, uint32_t cu_count) {
    if (a.full) {  // candidate collection: one gather shape per row length
        switch (chunks_per_lane) {
        case 1: return launch_t<1, 8, true>(a, waves_per_block, cu_count, s);
        case 2: return launch_t<2, 4, true>(a, waves_per_block, cu_count, s);
        case 3: return launch_t<3, 2, true>(a, waves_per_block, cu_count, s);
        case 4: return launch_t<4, 4, true>(a, waves_per_block, cu_count, s);
        case 8: return launch_t<8, 2, true>(a, waves_per_block, cu_count, s);
        case 16: return launch_t<16, 1, true>(a, waves_per_block, cu_count, s);
        default: return hipErrorInvalidValue;
        }
    }
    switch (chunks_per_lane) {
    case 1:
        return row_groups >= 16 ? launch_t<1, 16, false>(a, waves_per_block, cu_count, s)
                                : launch_t<1, 8, false>(a, waves_per_block, cu_count, s);
    case 2:
        return row_groups >= 16 ? launch_t<2, 16, false>(a, waves_per_block, cu_count, s)
             : row_groups >= 8 ? launch_t<2, 8, false>(a, waves_per_block, cu_count, s)
                               : launch_t<2, 4, false>(a, waves_per_block, cu_count, s);
    case 3:  // 8 lanes a row: G groups of 8 rows, the same rows per batch as 2 * G groups of 4
        return row_groups >= 16 ? launch_t<3, 8, false>(a, waves_per_block, cu_count, s)
             : row_groups >= 8 ? launch_t<3, 4, false>(a, waves_per_block, cu_count, s)
                               : launch_t<3, 2, false>(a, waves_per_block, cu_count, s);
    case 4:
        return row_groups >= 8 ? launch_t<4, 8, false>(a, waves_per_block, cu_count, s)
                               : launch_t<4, 4, false>(a, waves_per_block, cu_count, s);
    case 8:
        return launch_t<8, 2, false>(a, waves_per_block, cu_count, s);
    case 16:
        return launch_t<16, 1, false>(a, waves_per_block, cu_count, s);
    default:
        return hipErrorInvalidValue;
    }
}

}  // namespace wv
