// wv_search.hip -- greedy best-first search, one wavefront per query.
//
// Replaces ComponentSearchEntryRand::SearchEntryInner
// (src/component_search_entry.cpp:9-30) followed by
// ComponentSearchRouteGreedy::RouteInner (src/component_route.cpp:9-53),
// with Index::InsertIntoPool (include/weavess/index.h:152-180) done as a
// wave-wide merge.
//
// Per wavefront, in LDS: the sorted pool of L (dist,id) keys, its
// "unexpanded" flags, a 64-entry candidate block and an open-addressing
// visited table.  One expansion:
//   1. ballot over the flags finds the first unexpanded pool slot >= k;
//   2. lane i takes neighbour i of that node (64 per block; more in blocks);
//   3. LDS atomicCAS test-and-set in the visited table; survivors are
//      compacted;
//   4. distances, 16 lanes per row with float4 gathers, in the build order;
//   5. candidates not better than the pool's worst, or already in the pool,
//      are dropped; the rest are bitonic-sorted across the wave and merged
//      into the pool in place: a candidate's slot is its rank plus the number
//      of smaller pool keys, a pool key moves right by the number of smaller
//      candidates; keys are moved in 64-slot chunks from the tail backwards
//      (all reads of a chunk precede its writes), so slots before the first
//      insertion point are never touched;
//   6. the k-pointer restart of route.cpp:44-45 uses the final slot of the
//      best merged candidate, which equals the minimum insert position of
//      the sequential loop when (dist,id) is a strict order.
// The visited table forgets: when an id finds no free slot within the probe
// limit it evicts the occupant of its home slot.  A forgotten node that comes
// back is re-evaluated; it is then either already in the pool (duplicate) or
// no better than the pool's worst, so ids and distances are unchanged and
// only the evaluation count grows.
#include "wv_device.hpp"
#include "wv_internal.hpp"
#include "wv_mlp.hpp"

namespace wv {

namespace {

constexpr int VIS_PROBES = 8;

// Test-and-set of `id` in the visited table; true when this call inserted it
// (first visit, or the id had been forgotten: when no free slot is found
// within the probe limit the occupant of the home slot is evicted).
__device__ inline bool visit_first(uint32_t *vis, uint32_t hmask, uint32_t id) {
    const uint32_t h0 = fmix32(id) & hmask;
    uint32_t h = h0;
    for (int p = 0; p < VIS_PROBES; ++p) {
        const uint32_t old = atomicCAS(&vis[h], ID_NONE, id);
        if (old == ID_NONE) return true;
        if (old == id) return false;
        h = (h + 1) & hmask;
    }
    atomicExch(&vis[h0], id);
    return true;
}

// SearchEntryRand's GenRandom (include/weavess/util.h:13-27) on the query's
// counter stream: L draws mod (N-L), sorted, made strictly increasing
// (a'[i] = i + max_{j<=i}(a[j]-j), a prefix-max scan), rotated by one more
// draw mod N.  a has P >= L slots.
__device__ inline void gen_entries(uint64_t seed, uint32_t q, uint32_t L, uint32_t N, uint32_t P,
                                   uint32_t *a) {
    const int lane = lane_id();
    const uint64_t s = stream_of(seed, TAG_ENTRY, q);
    const uint32_t mod = N - L;
    for (uint32_t i = lane; i < P; i += WAVE) a[i] = i < L ? draw(s, i) % mod : 0xFFFFFFFFu;
    wave_sync();
    lds_sort32(a, (int)P);
    const int seg = (int)((L + WAVE - 1) / WAVE);
    const int s0 = lane * seg, s1 = min(s0 + seg, (int)L);
    int64_t mx = INT64_MIN;
    for (int i = s0; i < s1; ++i) {
        int64_t v = (int64_t)a[i] - i;
        if (v > mx) mx = v;
    }
    int64_t incl = wave_scan_max64(mx);
    uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)incl, 1);
    uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)((uint64_t)incl >> 32), 1);
    int64_t run = lane == 0 ? INT64_MIN : (int64_t)(((uint64_t)hi << 32) | lo);
    const uint64_t off = draw(s, L) % N;
    for (int i = s0; i < s1; ++i) {
        int64_t v = (int64_t)a[i] - i;
        if (v > run) run = v;
        a[i] = (uint32_t)(((uint64_t)(run + i) + off) % N);
    }
    wave_sync();
}

template <int CPL, int G>
__global__ void __launch_bounds__(256) greedy_search_kernel(SearchArgs a) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int wid = (int)(threadIdx.x >> 6);
    const int lane = lane_id();
    const uint32_t q = blockIdx.x * (blockDim.x >> 6) + wid;
    if (q >= a.Q) return;

    uint8_t *my = smem + (size_t)wid * a.wave_lds;
    uint64_t *pool = reinterpret_cast<uint64_t *>(my);         // P keys
    uint64_t *ckey = pool + a.P;                                // 64
    uint32_t *cid = reinterpret_cast<uint32_t *>(ckey + WAVE);  // 64
    uint32_t *vis = cid + WAVE;                                 // H (>= P)
    uint8_t *flag = reinterpret_cast<uint8_t *>(vis + a.H);     // P

    const uint32_t L = a.L_per_q ? a.L_per_q[q] : a.L;
    const uint32_t kq = a.k_per_q ? a.k_per_q[q] : a.k;
    const uint32_t N = a.N, nch = a.nch, dp = a.dp;

    float4 qv[CPL];
    {
        const float4 *qp = reinterpret_cast<const float4 *>(a.queries + (size_t)q * dp);
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            const int ch = (lane & 15) + 16 * c;
            qv[c] = ch < (int)nch ? qp[ch] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }

    // ---- entries (SearchEntryRand): not marked visited, all flagged; the
    // visited table's storage holds the entry ids until they are scored
    uint32_t *ent = vis;
    if (a.entries) {
        for (uint32_t i = lane; i < L; i += WAVE) ent[i] = a.entries[(size_t)q * L + i];
        wave_sync();
    } else {
        gen_entries(a.seed, q, L, N, a.P, ent);
    }
    rows_dist<CPL, G>(a.base, dp, nch, qv, ent, (int)L, pool);
    for (uint32_t i = L + lane; i < a.P; i += WAVE) pool[i] = KEY_EMPTY;
    wave_sync();
    lds_sort64(pool, (int)a.P);
    const uint32_t hmask = a.H - 1;
    for (uint32_t i = lane; i < a.H; i += WAVE) vis[i] = ID_NONE;
    for (uint32_t i = lane; i < a.P; i += WAVE) flag[i] = i < L ? 1 : 0;
    wave_sync();

    uint32_t evals = L, hops = 0;
    int k = 0;
    const int Li = (int)L;
    // learned termination / traces: state features (identical on every lane)
    const bool watch = a.model != nullptr || a.trace != nullptr;
    const int kidx = (int)(kq < L ? kq : L) - 1;
    const float entry_d = watch ? key_dist(pool[0]) : 0.f;
    // k-th distance 1..4 expansions ago
    float r0 = watch ? key_dist(pool[kidx]) : 0.f, r1 = r0, r2 = r0, r3 = r0;
    uint32_t settle = 0;  // last expansion that changed the top-k
    for (;;) {
        // first unexpanded slot >= k
        int kk = Li;
        for (int t0 = (k & ~(WAVE - 1)); t0 < Li; t0 += WAVE) {
            const int p = t0 + lane;
            const bool f = p >= k && p < Li && flag[p];
            const uint64_t b = __ballot(f);
            if (b) {
                kk = t0 + __ffsll((unsigned long long)b) - 1;
                break;
            }
        }
        if (watch && hops > 0) {
            // state after `hops` expansions
            float f[MLP_IN];
            const float kth = key_dist(pool[kidx]);
            term_features((float)hops, kth, key_dist(pool[0]), entry_d, r3, (float)(hops - settle), (float)L,
                          (float)kq, f);
            r3 = r2;
            r2 = r1;
            r1 = r0;
            r0 = kth;
            if (a.trace && hops <= a.trace_cap && lane < MLP_IN)
                a.trace[((size_t)q * a.trace_cap + (hops - 1)) * MLP_IN + lane] = pick8(f, lane);
            if (kk < Li && a.model && hops >= a.model->min_hops && term_logit(a.model, f) > a.model->logit_threshold)
                break;
        }
        if (kk >= Li) break;
        const uint32_t u = key_id(pool[kk]);
        wave_sync();
        if (lane == 0) flag[kk] = 0;
        ++hops;
        uint32_t du = a.deg[u];
        if (du > a.deg_max) du = a.deg_max;
        int nk = Li;
        const uint32_t *adj = a.graph + (size_t)u * a.deg_max;
        for (uint32_t m0 = 0; m0 < du; m0 += WAVE) {
            const uint32_t mi = m0 + lane;
            const uint32_t id = mi < du ? adj[mi] : ID_NONE;
            const bool fresh = id < N && visit_first(vis, hmask, id);
            const uint64_t bm = __ballot(fresh);
            const int m = __popcll(bm);
            if (fresh) cid[lanes_below(bm)] = id;
            wave_sync();
            if (m == 0) continue;
            rows_dist<CPL, G>(a.base, dp, nch, qv, cid, m, ckey);
            evals += (uint32_t)m;
            wave_sync();
            const uint64_t worst = pool[Li - 1];
            uint64_t key = lane < m ? ckey[lane] : KEY_EMPTY;
            int lb = Li;
            if (key < worst) {
                lb = lower_bound64(pool, Li, key);
                if (pool[lb] == key) key = KEY_EMPTY;  // already pooled: duplicate
            } else {
                key = KEY_EMPTY;
            }
            if (__ballot(key != KEY_EMPTY) == 0) continue;
            // sort (key, lb) pairs by key: lb is monotone in key, recompute after
            key = wave_sort64(key);
            const uint64_t prev = shfl64(key, lane > 0 ? lane - 1 : 0);
            const bool live = key != KEY_EMPTY && !(lane > 0 && prev == key);
            const uint64_t lm = __ballot(live);
            const int c = __popcll(lm);
            const int r = (int)lanes_below(lm);
            wave_sync();
            if (live) ckey[r] = key;
            wave_sync();
            // final slot of each live candidate, from the unmoved pool
            int cpos = Li;
            if (live) cpos = r + lower_bound64(pool, Li, key);
            const int pos0 = lower_bound64(pool, Li, ckey[0]);
            // shift pool keys right by the number of smaller candidates,
            // 64-slot chunks from the tail back to the chunk holding pos0
            for (int t0 = (Li - 1) & ~(WAVE - 1); t0 >= (pos0 & ~(WAVE - 1)); t0 -= WAVE) {
                const int p = t0 + lane;
                const bool mv = p >= pos0 && p < Li;
                uint64_t pk = 0;
                uint8_t pf = 0;
                int np = Li;
                if (mv) {
                    pk = pool[p];
                    pf = flag[p];
                    np = p + lower_bound64(ckey, c, pk);
                }
                wave_sync();
                if (mv && np < Li) {
                    pool[np] = pk;
                    flag[np] = pf;
                }
                wave_sync();
            }
            if (live && cpos < Li) {
                pool[cpos] = key;
                flag[cpos] = 1;
            }
            if (pos0 < nk) nk = pos0;
            if (pos0 <= kidx) settle = hops;  // a candidate entered the top-k
            wave_sync();
        }
        k = (nk <= kk) ? nk : kk + 1;
    }

    for (uint32_t i = lane; i < a.k_stride; i += WAVE) {
        const bool ok = i < kq;
        a.out_ids[(size_t)q * a.k_stride + i] = ok ? key_id(pool[i]) : ID_NONE;
        if (a.out_dist) a.out_dist[(size_t)q * a.k_stride + i] = ok ? key_dist(pool[i]) : -1.0f;
    }
    if (lane == 0) {
        if (a.evals) a.evals[q] = evals;
        if (a.hops) a.hops[q] = hops;
        if (a.trace_len) a.trace_len[q] = hops < a.trace_cap ? hops : a.trace_cap;
        if (a.settle) a.settle[q] = settle;
    }
}

template <int CPL, int G>
hipError_t launch_t(const SearchArgs &a, int waves_per_block, hipStream_t s) {
    const size_t lds = (size_t)a.wave_lds * waves_per_block;
    const uint32_t blocks = (a.Q + waves_per_block - 1) / waves_per_block;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&greedy_search_kernel<CPL, G>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((greedy_search_kernel<CPL, G>), dim3(blocks), dim3(WAVE * waves_per_block),
                       lds, s, a);
    return hipGetLastError();
}

}  // namespace

size_t search_wave_lds(uint32_t P, uint32_t H) {
    size_t b = (size_t)P * 8 + WAVE * 8 + WAVE * 4 + (size_t)H * 4 + (size_t)P;
    return (b + 15) & ~(size_t)15;
}

hipError_t launch_greedy_search(const SearchArgs &a, int chunks_per_lane, int row_groups,
                                int waves_per_block, hipStream_t s) {
    switch (chunks_per_lane) {
    case 1:
        return row_groups >= 16 ? launch_t<1, 16>(a, waves_per_block, s) : launch_t<1, 8>(a, waves_per_block, s);
    case 2:
        return row_groups >= 16 ? launch_t<2, 16>(a, waves_per_block, s)
             : row_groups >= 8 ? launch_t<2, 8>(a, waves_per_block, s)
                               : launch_t<2, 4>(a, waves_per_block, s);
    case 4:
        return row_groups >= 8 ? launch_t<4, 8>(a, waves_per_block, s) : launch_t<4, 4>(a, waves_per_block, s);
    case 8:
        return launch_t<8, 2>(a, waves_per_block, s);
    case 16:
        return launch_t<16, 1>(a, waves_per_block, s);
    default:
        return hipErrorInvalidValue;
    }
}

}  // namespace wv
