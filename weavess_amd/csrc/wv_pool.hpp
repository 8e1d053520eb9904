// wv_pool.hpp -- sorted key arrays in LDS shared by the greedy search and
// the NSG refine: entry keys with an "unexpanded" flag, binary searches, the
// in-place merge of a sorted block, and GenRandom into LDS.
#pragma once

#include "wv_device.hpp"

namespace wv {

// Pool and buffer entries are keys with bit 31 (the top bit of the id word;
// ids are below 2^31) set while the entry is unexpanded.  Comparisons use the
// key with that bit cleared, so the (dist, id) order is unchanged, and a merge
// moves the flag with its key.
constexpr uint64_t FBIT = 1ull << 31;
__device__ inline uint64_t canon(uint64_t e) { return e & ~FBIT; }
__device__ inline uint32_t entry_id(uint64_t e) { return (uint32_t)e & 0x7FFFFFFFu; }

// Number of entries of the sorted a[0..n) whose key is < key (<= key).
__device__ inline int lower_bound_e(const uint64_t *a, int n, uint64_t key) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (canon(a[mid]) < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}
__device__ inline int upper_bound_e(const uint64_t *a, int n, uint64_t key) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (canon(a[mid]) <= key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// One round of an 8-ary search for the number of entries < key in a sorted
// range: the answer lies in [lo, hi] (entries before lo are < key, from hi
// on >= key) and hi - lo >= 8.  The 7 slots lo + j*s (s = (hi-lo)/8, j =
// 1..7) are read at once -- independent LDS reads instead of a chain of
// dependent ones -- and the range narrows to the part between the last slot
// below key and the first one not below it (at most s + 6 slots).
__device__ inline void narrow8_e(const uint64_t *a, int &lo, int &hi, uint64_t key) {
    const int s = (hi - lo) >> 3;
    uint64_t x[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) x[j] = canon(a[lo + (j + 1) * s]);
    int t = 0;
#pragma unroll
    for (int j = 0; j < 7; ++j) t += x[j] < key ? 1 : 0;
    const int nlo = t > 0 ? lo + t * s + 1 : lo;
    hi = t < 7 ? lo + (t + 1) * s : hi;
    lo = nlo;
}

// The last round for a range of at most 7 slots: all of them read at once;
// returns lo plus the number below key.
__device__ inline int finish8_e(const uint64_t *a, int lo, int hi, uint64_t key) {
    uint64_t x[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) x[j] = lo + j < hi ? canon(a[lo + j]) : KEY_EMPTY;
    int t = 0;
#pragma unroll
    for (int j = 0; j < 7; ++j) t += x[j] < key ? 1 : 0;
    return lo + t;
}

// Number of entries of the sorted a[0..n) whose key is < key: 8-ary rounds
// down to at most 7 slots, then one round over those.  The result is
// lower_bound_e's; only the reads are fewer rounds deep.
__device__ inline int lower_bound8_e(const uint64_t *a, int n, uint64_t key) {
    int lo = 0, hi = n;
    while (hi - lo >= 8) narrow8_e(a, lo, hi, key);
    return finish8_e(a, lo, hi, key);
}

// Merge c sorted entries src[0..c) into the sorted array arr[0..len) of
// capacity cap, in place: an entry of arr moves right by the number of
// smaller src keys, src[j] lands at j + the number of arr keys <= it (so
// equal keys never collide), and entries pushed to slots >= cap are dropped.
// With `mark`, the src entries are stored unexpanded (bit 31 set); without,
// as they are.  Entries are moved in 64-slot chunks from the tail backwards
// (all reads of a chunk precede its writes), so slots before the first
// insertion point are never touched.  1 <= c <= 4 * WAVE.  Returns the final
// slot of src[0] (>= cap when it was dropped).
__device__ inline int merge_into(uint64_t *arr, int len, int cap, const uint64_t *src, int c, bool mark) {
    const int lane = lane_id();
    int cpos[4], ub[4];
    uint64_t ck[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int j = lane + r * WAVE;
        cpos[r] = cap;
        ub[r] = 0x7FFFFFFF;
        ck[r] = 0;
        if (j < c) {
            ck[r] = mark ? (src[j] | FBIT) : src[j];
            ub[r] = lower_bound8_e(arr, len, canon(ck[r]) + 1);  // arr keys <= src[j]
            cpos[r] = j + ub[r];
        }
    }
    // The src keys below arr[p] are those whose ub (non-decreasing in j) is
    // <= p.  For a chunk [t0, t0 + 64): the ub below t0 all count (a ballot),
    // and the few in the chunk are read from their lanes one by one.
    const int pos0 = __shfl(cpos[0], 0);  // lane 0 holds src[0]: j = 0
    for (int t0 = (len - 1) & ~(WAVE - 1); t0 >= (pos0 & ~(WAVE - 1)); t0 -= WAVE) {
        const int p = t0 + lane;
        const bool mv = p >= pos0 && p < len;
        uint64_t pk = 0;
        if (mv) pk = arr[p];
        int base = 0, top = 0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            if (r * WAVE >= c) break;
            base += (int)__popcll(__ballot(ub[r] < t0));
            top += (int)__popcll(__ballot(ub[r] < t0 + WAVE));
        }
        int below = base;
        for (int jj = base; jj < top; ++jj) {
            const int rr = jj / WAVE;
            const int u = rr == 0 ? ub[0] : rr == 1 ? ub[1] : rr == 2 ? ub[2] : ub[3];
            below += __builtin_amdgcn_readlane(u, jj % WAVE) <= p ? 1 : 0;
        }
        const int np = p + below;
        wave_sync();
        if (mv && np < cap) arr[np] = pk;
        wave_sync();
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
        if (cpos[r] < cap) arr[cpos[r]] = ck[r];
    wave_sync();
    return pos0;
}

// merge_into(buf, nb, B, src, c, true) for 1 <= c <= 64 sorted, distinct
// keys that are not in buf (lane t < c holds src[t] in `skey`, without the
// flag bit), nb + c <= B <= 256, followed by the new L-th key of pool +
// buffer (kth_union(pool, Li, buf, nb + c, Li - 1)) into `worst`.  Both come
// from registers and ballots instead of chains of dependent LDS searches:
// every lane reads its buffer entries (l, l + 64, ...); a loop over the
// src keys counts, per entry, the src keys below it (its shift right) and,
// per src key, the entries below it (a ballot); all keys are then written to
// their final slots at once.  New buffer entry j is among the L smallest of
// pool + buffer iff it is below pool[L-1-j] (true for a prefix of j), so
// the number i taken from the buffer is a ballot count and the L-th key is
// the larger of buffer entry i-1 and pool entry L-1-i.  Returns the final
// slot of src[0].
__device__ inline int buffer_insert(uint64_t *buf, int nb, uint64_t skey, int c, const uint64_t *pool, int Li,
                                    uint64_t &worst) {
    const int lane = lane_id();
    constexpr int R = 4;  // entries a lane holds: B <= 256
    const int nn = nb + c, hi = nn < Li ? nn : Li;
    uint64_t b[R], pt[R];
    int np[R];
    // the pool's last entries (read first: the merge below does not change
    // the pool, and their LDS reads then travel with the buffer's)
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = lane + r * WAVE;
        b[r] = j < nb ? buf[j] : KEY_EMPTY;
        pt[r] = j < Li && r * WAVE <= hi ? canon(pool[Li - 1 - j]) : 0;  // entry taken (<= hi) may be read
        np[r] = j;
    }
    int spos = 0;
    for (int t = 0; t < c; ++t) {
        const uint64_t s = readlane64(skey, t);
        int below = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r * WAVE >= nb) break;
            const bool in = lane + r * WAVE < nb;
            const uint64_t cb = canon(b[r]);
            if (in && s < cb) ++np[r];
            below += (int)__popcll(__ballot(in && cb < s));
        }
        if (lane == t) spos = t + below;
    }
    wave_sync();
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (lane + r * WAVE < nb) buf[np[r]] = b[r];
    if (lane < c) buf[spos] = skey | FBIT;
    wave_sync();
    uint64_t bn[R];
    int taken = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = lane + r * WAVE;
        bn[r] = j < nn ? canon(buf[j]) : KEY_EMPTY;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) taken += (int)__popcll(__ballot(lane + r * WAVE < hi && bn[r] < pt[r]));
    uint64_t x = 0, y = 0;
    if (taken > 0) {
        const int i = taken - 1, rr = i / WAVE;
        x = readlane64(rr == 0 ? bn[0] : rr == 1 ? bn[1] : rr == 2 ? bn[2] : bn[3], i % WAVE);
    }
    if (taken < Li) {
        const int rr = taken / WAVE;
        y = rr < R ? readlane64(rr == 0 ? pt[0] : rr == 1 ? pt[1] : rr == 2 ? pt[2] : pt[3], taken % WAVE)
                   : canon(pool[Li - 1 - taken]);
    }
    worst = x > y ? x : y;
    return __shfl(spos, 0);
}

// GenRandom (include/weavess/util.h:13-27) of size L on counter stream s
// (SearchEntryRand's entries use the query's stream): L draws mod (N-L), sorted, made strictly increasing
// (a'[i] = i + max_{j<=i}(a[j]-j), a prefix-max scan), rotated by one more
// draw mod N.  The draws are sorted as (value, index) keys, 64 at a time,
// merged into a[0..L) (index i < 2^31, so bit 31 stays clear); on return
// a[i] holds entry i's id.  tmp: 64 keys of scratch.
__device__ inline void gen_random_lds(uint64_t s, uint32_t L, uint32_t N, uint64_t *a, uint64_t *tmp) {
    const int lane = lane_id();
    const uint32_t mod = N - L;
    int len = 0;
    for (uint32_t c0 = 0; c0 < L; c0 += WAVE) {
        const uint32_t i = c0 + lane;
        const uint64_t key = wave_sort64(i < L ? ((uint64_t)(draw(s, i) % mod) << 32) | i : KEY_EMPTY);
        const int cnt = (int)min((uint32_t)WAVE, L - c0);
        tmp[lane] = key;
        wave_sync();
        merge_into(a, len, (int)L, tmp, cnt, false);
        len += cnt;
    }
    const int seg = (int)((L + WAVE - 1) / WAVE);
    const int s0 = lane * seg, s1 = min(s0 + seg, (int)L);
    int64_t mx = INT64_MIN;
    for (int i = s0; i < s1; ++i) {
        int64_t v = (int64_t)(a[i] >> 32) - i;
        if (v > mx) mx = v;
    }
    int64_t incl = wave_scan_max64(mx);
    uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)incl, 1);
    uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)((uint64_t)incl >> 32), 1);
    int64_t run = lane == 0 ? INT64_MIN : (int64_t)(((uint64_t)hi << 32) | lo);
    const uint64_t off = draw(s, L) % N;
    for (int i = s0; i < s1; ++i) {
        int64_t v = (int64_t)(a[i] >> 32) - i;
        if (v > run) run = v;
        a[i] = ((uint64_t)(run + i) + off) % N;
    }
    wave_sync();
}

// Merge the distinct keys blk[0..m) (m <= 64, any order, no flag bit) into
// the sorted distinct array full[0..len) of capacity cap, dropping keys
// already there and keys that cannot enter a full array.  tmp: 64 keys of
// scratch.  Returns the new length.
__device__ inline int merge_distinct(uint64_t *full, int len, int cap, const uint64_t *blk, int m, uint64_t *tmp) {
    const int lane = lane_id();
    uint64_t k = lane < m ? blk[lane] : KEY_EMPTY;
    if (k != KEY_EMPTY) {
        if (len == cap && k >= full[cap - 1]) {
            k = KEY_EMPTY;
        } else {
            const int p = lower_bound_e(full, len, k);
            if (p < len && full[p] == k) k = KEY_EMPTY;
        }
    }
    const uint64_t lm = __ballot(k != KEY_EMPTY);
    if (lm == 0) return len;
    const int c = __popcll(lm);
    if (k != KEY_EMPTY) tmp[lanes_below(lm)] = k;
    wave_sync();
    int w = 2;
    while (w < c) w <<= 1;
    k = wave_sort64_n(lane < c ? tmp[lane] : KEY_EMPTY, w);
    wave_sync();
    if (lane < c) tmp[lane] = k;
    wave_sync();
    merge_into(full, len, cap, tmp, c, false);
    return min(len + c, cap);
}

// Collection staged in blocks of 64: keys of `blk` (m <= 64) that can still
// enter the sorted distinct top-cap list `full` wait in `stg` (128 keys, sn
// of them); every 64 staged keys are merged at once.  A key at or above a
// full list's last entry is dropped on arrival: that entry only falls.  The
// final list is the same as merging every block as it comes (the cap
// smallest distinct keys seen), for fewer passes over `full`.
__device__ inline int merge_staged(uint64_t *full, int len, int cap, const uint64_t *stg, int m, uint64_t *tmp) {
    const int lane = lane_id();
    // sorted, then repeated keys (one id evaluated twice) dropped
    uint64_t k = wave_sort64(lane < m ? stg[lane] : KEY_EMPTY);
    const uint64_t prev = shfl64(k, lane > 0 ? lane - 1 : 0);
    if (lane > 0 && prev == k) k = KEY_EMPTY;
    if (k != KEY_EMPTY) {
        if (len == cap && k >= full[cap - 1]) {
            k = KEY_EMPTY;
        } else {
            const int p = lower_bound_e(full, len, k);
            if (p < len && full[p] == k) k = KEY_EMPTY;
        }
    }
    const uint64_t lm = __ballot(k != KEY_EMPTY);
    if (lm == 0) return len;
    const int c = __popcll(lm);
    wave_sync();
    if (k != KEY_EMPTY) tmp[lanes_below(lm)] = k;  // still ascending
    wave_sync();
    merge_into(full, len, cap, tmp, c, false);
    return min(len + c, cap);
}
__device__ inline void collect_stage(uint64_t *full, int &len, int cap, const uint64_t *blk, int m, uint64_t *stg,
                                     int &sn, uint64_t *tmp) {
    const int lane = lane_id();
    uint64_t k = lane < m ? blk[lane] : KEY_EMPTY;
    if (k != KEY_EMPTY && len == cap && k >= full[cap - 1]) k = KEY_EMPTY;
    const uint64_t b = __ballot(k != KEY_EMPTY);
    if (k != KEY_EMPTY) stg[sn + (int)lanes_below(b)] = k;
    sn += __popcll(b);
    if (sn >= WAVE) {
        wave_sync();
        len = merge_staged(full, len, cap, stg, WAVE, tmp);
        const uint64_t r = lane + WAVE < sn ? stg[lane + WAVE] : KEY_EMPTY;
        wave_sync();
        if (lane + WAVE < sn) stg[lane] = r;
        sn -= WAVE;
        wave_sync();
    }
}
__device__ inline int collect_flush(uint64_t *full, int len, int cap, const uint64_t *stg, int sn, uint64_t *tmp) {
    wave_sync();
    return sn > 0 ? merge_staged(full, len, cap, stg, sn, tmp) : len;
}

}  // namespace wv
