// wv_device.hpp -- device-side building blocks shared by the HIP kernels.
//
// Everything here is written for gfx950 (CDNA4): 64-lane wavefronts, LDS
// atomics, 64-bit keys moved across lanes as two 32-bit shuffles.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wv {

constexpr int WAVE = 64;
constexpr size_t LDS_PER_CU = 160 * 1024;  // gfx950: LDS bytes a compute unit holds
constexpr uint64_t KEY_EMPTY = ~0ull;
constexpr uint32_t ID_NONE = 0xFFFFFFFFu;
constexpr uint64_t GOLDEN = 0x9E3779B97F4A7C15ull;

// Stream tags; the values are part of the seeded-output contract.
constexpr uint64_t TAG_INIT = 0x494E4954ull;    // "INIT"
constexpr uint64_t TAG_ENTRY = 0x454E5452ull;   // "ENTR"
constexpr uint64_t TAG_REVNEW = 0x52564E57ull;  // "RVNW"
constexpr uint64_t TAG_REVOLD = 0x52564F4Cull;  // "RVOL"
constexpr uint64_t TAG_EFANNA = 0x45464E41ull;  // "EFNA": EFANNA's first-round samples

// ---------------------------------------------------------------- keys
// (dist, id) packed so that unsigned 64-bit order is (dist, id) order for
// non-negative finite distances; ties between equal distances go to the
// smaller id.
__host__ __device__ inline uint64_t make_key(float d, uint32_t id) {
    uint32_t b;
    __builtin_memcpy(&b, &d, 4);
    return ((uint64_t)b << 32) | id;
}
__host__ __device__ inline uint32_t key_id(uint64_t k) { return (uint32_t)k; }
__host__ __device__ inline float key_dist(uint64_t k) {
    const uint32_t b = (uint32_t)(k >> 32);
    float f;
    __builtin_memcpy(&f, &b, 4);
    return f;
}

// ---------------------------------------------------------------- hashing
// splitmix64 finaliser (Steele, Lea, Flood 2014) and murmur3's fmix32.
__host__ __device__ inline uint64_t fmix64(uint64_t z) {
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    return z;
}
__host__ __device__ inline uint32_t fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}
// Counter-based random stream: draw k of stream s is independent of every
// other draw, so a wavefront can produce draws in any order.
__host__ __device__ inline uint64_t stream_of(uint64_t seed, uint64_t tag, uint64_t idx) {
    uint64_t h = fmix64(seed + GOLDEN);
    h = fmix64(h ^ (tag * GOLDEN));
    h = fmix64(h ^ (idx + 0x632BE59BD9B4E019ull));
    return h;
}
__host__ __device__ inline uint32_t draw(uint64_t s, uint64_t k) {
    return (uint32_t)(fmix64(s + (k + 1) * GOLDEN) >> 32);
}
// Slot of candidate `id` in target `t`'s candidate table.
__host__ __device__ inline uint32_t cand_slot(uint32_t t, uint32_t id, uint32_t mask) {
    return fmix32(id * 0x9E3779B1u + t * 0x7F4A7C15u) & mask;
}

// ---------------------------------------------------------------- lanes
__device__ inline int lane_id() { return (int)(threadIdx.x & (WAVE - 1)); }

// Make this wave's LDS writes visible to its own lanes (no cross-wave
// barrier: waves of one block run independent queries/nodes).
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// v of lane (this lane XOR m).  m = 1, 2, 4, 8 move within a row of 16
// lanes by DPP (no LDS round trip): quad_perm [1,0,3,2] and [2,3,0,1];
// row_half_mirror then quad_perm [3,2,1,0] (lane i of 8 takes 7 - (i ^ 3)
// = i ^ 4); row_ror:8.  Other m go through ds_bpermute.  Every lane must be
// active (a DPP read of an inactive lane is not its register).
__device__ inline uint32_t xor_lane(uint32_t v, int m) {
    const int x = (int)v;
    if (m == 1) return (uint32_t)__builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false);
    if (m == 2) return (uint32_t)__builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false);
    if (m == 4)
        return (uint32_t)__builtin_amdgcn_update_dpp(0, __builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, false),
                                                     0x1B, 0xF, 0xF, false);
    if (m == 8) return (uint32_t)__builtin_amdgcn_update_dpp(0, x, 0x128, 0xF, 0xF, false);
    return (uint32_t)__shfl_xor(x, m);
}
__device__ inline float xor_lane(float v, int m) { return __uint_as_float(xor_lane(__float_as_uint(v), m)); }
__device__ inline uint64_t xor_lane64(uint64_t v, int m) {
    return ((uint64_t)xor_lane((uint32_t)(v >> 32), m) << 32) | xor_lane((uint32_t)v, m);
}

__device__ inline uint64_t shfl_xor64(uint64_t v, int m) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    lo = (uint32_t)__shfl_xor((int)lo, m);
    hi = (uint32_t)__shfl_xor((int)hi, m);
    return ((uint64_t)hi << 32) | lo;
}
__device__ inline uint64_t shfl64(uint64_t v, int src) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    lo = (uint32_t)__shfl((int)lo, src);
    hi = (uint32_t)__shfl((int)hi, src);
    return ((uint64_t)hi << 32) | lo;
}

// Lane l's value of a 64-bit register, l the same on every lane.
__device__ inline uint64_t readlane64(uint64_t v, int l) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

// Number of set bits of `mask` below this lane.
__device__ inline uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Ascending bitonic sort of one 64-bit key per lane across the wavefront
// (every lane active, as for all wavefront-wide exchanges here).  DPP:
// exchanges within 16 lanes by DPP (xor_lane) instead of ds_bpermute,
// which shortens a latency-bound wavefront's chain but costs VALU issue
// where the SIMD is busy.
template <bool DPP = false>
__device__ inline uint64_t wave_sort64(uint64_t key) {
    const int lane = lane_id();
#pragma unroll
    for (int k = 2; k <= WAVE; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            uint64_t o = DPP ? xor_lane64(key, j) : shfl_xor64(key, j);
            bool up = (lane & k) == 0;
            bool lower = (lane & j) == 0;
            bool keep_min = (lower == up);
            uint64_t mn = key < o ? key : o, mx = key < o ? o : key;
            key = keep_min ? mn : mx;
        }
    }
    return key;
}

// Ascending bitonic sort of the keys in lanes [0, w), w a power of two in
// [2, 64] (lanes from w up keep KEY_EMPTY-sorted garbage of their own block).
template <bool DPP = false>
__device__ inline uint64_t wave_sort64_n(uint64_t key, int w) {
    const int lane = lane_id();
    for (int k = 2; k <= w; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            uint64_t o = DPP ? xor_lane64(key, j) : shfl_xor64(key, j);
            bool up = (lane & k) == 0;
            bool lower = (lane & j) == 0;
            bool keep_min = (lower == up);
            uint64_t mn = key < o ? key : o, mx = key < o ? o : key;
            key = keep_min ? mn : mx;
        }
    }
    return key;
}

// Ascending bitonic sort of P (power of two, >= 2) 64-bit keys in LDS by one
// wavefront.
__device__ inline void lds_sort64(uint64_t *a, int P) {
    const int lane = lane_id();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = lane; t < P / 2; t += WAVE) {
                int lo = ((t & ~(j - 1)) << 1) | (t & (j - 1));
                int hi = lo | j;
                bool up = (lo & k) == 0;
                uint64_t x = a[lo], y = a[hi];
                if ((x > y) == up) {
                    a[lo] = y;
                    a[hi] = x;
                }
            }
            wave_sync();
        }
    }
}
__device__ inline void lds_sort32(uint32_t *a, int P) {
    const int lane = lane_id();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = lane; t < P / 2; t += WAVE) {
                int lo = ((t & ~(j - 1)) << 1) | (t & (j - 1));
                int hi = lo | j;
                bool up = (lo & k) == 0;
                uint32_t x = a[lo], y = a[hi];
                if ((x > y) == up) {
                    a[lo] = y;
                    a[hi] = x;
                }
            }
            wave_sync();
        }
    }
}

// Number of elements of sorted a[0..n) that are < key.
__device__ inline int lower_bound64(const uint64_t *a, int n, uint64_t key) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Number of keys <= key in the sorted a[0..n).
__device__ inline int upper_bound64(const uint64_t *a, int n, uint64_t key) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] <= key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Inclusive wave-wide max scan of a signed 64-bit value.
__device__ inline int64_t wave_scan_max64(int64_t v) {
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)v, o);
        uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)((uint64_t)v >> 32), o);
        int64_t u = (int64_t)(((uint64_t)hi << 32) | lo);
        if (lane >= o && u > v) v = u;
    }
    return v;
}

// ---------------------------------------------------------------- distance
// One chunk (four consecutive elements) of the build-order distance:
// ((x0^2 + x1^2) + x2^2) + x3^2 with no fused multiply-add (the library is
// compiled with -ffp-contract=off).
__device__ inline float chunk_sq(const float4 a, const float4 b) {
    float d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z, d3 = a.w - b.w;
    float s = d0 * d0 + d1 * d1;
    s = s + d2 * d2;
    s = s + d3 * d3;
    return s;
}
// Fold of the 16 accumulators of a 16-lane group: acc[j] += acc[j+h] for
// h = 8, 4, 2, 1.  Every lane of the group ends with the same value.
template <bool DPP = false>
__device__ inline float fold16(float acc) {
    acc = acc + (DPP ? xor_lane(acc, 8) : __shfl_xor(acc, 8));
    acc = acc + (DPP ? xor_lane(acc, 4) : __shfl_xor(acc, 4));
    acc = acc + (DPP ? xor_lane(acc, 2) : __shfl_xor(acc, 2));
    acc = acc + (DPP ? xor_lane(acc, 1) : __shfl_xor(acc, 1));
    return acc;
}

// Lanes that share one row in rows_dist: 16, lane j owning accumulator j of
// the build order (chunks j, j+16, ...); or, for rows of 17-24 chunks
// (CPL 3), 8, lane j owning accumulators j and j+8 (chunks j, j+8, j+16
// alternate between them), whose sum is the first step of the fold.  The
// second form reads 24 chunks of a row in 3 loads per lane instead of 32
// slots in 2 (half of the second masked off), with the same sums in the
// same order.
template <int CPL> constexpr int row_lanes() { return CPL == 3 ? 8 : 16; }

// The query's chunks in the lane order rows_dist expects.
template <int CPL>
__device__ inline void load_query(const float *q, uint32_t nch, float4 (&qv)[CPL]) {
    constexpr int LPR = row_lanes<CPL>();
    const int sub = lane_id() % LPR;
    const float4 *qp = reinterpret_cast<const float4 *>(q);
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
        const int ch = sub + LPR * c;
        qv[c] = ch < (int)nch ? qp[ch] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// rows_dist in two halves: rows_issue gathers the first G row groups (the
// rows a wavefront holds at once) into v; rows_finish reduces them and then
// gathers and reduces any further rows.  Between the two a caller can do
// other work while the loads travel.
template <int CPL, int G>
__device__ inline void rows_issue(const float *__restrict__ base, uint32_t dp, uint32_t nch, const uint32_t *ids,
                                  int m, int r0, float4 (&v)[G][CPL]) {
    constexpr int LPR = row_lanes<CPL>(), RPG = WAVE / LPR;
    const int lane = lane_id(), sub = lane % LPR, grp = lane / LPR;
    // groups wholly past m (the same for the whole wavefront) are skipped
#pragma unroll
    for (int g = 0; g < G; ++g) {
        if (r0 + RPG * g >= m) break;
        const int row = r0 + RPG * g + grp;
        const uint32_t id = row < m ? ids[row] : 0u;
        const float4 *p = reinterpret_cast<const float4 *>(base + (size_t)id * dp);
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            const int ch = sub + LPR * c;
            v[g][c] = (row < m && ch < (int)nch) ? p[ch] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
}

template <int CPL, int G, bool DPP = false>
__device__ inline void rows_reduce(uint32_t nch, const float4 (&qv)[CPL], const uint32_t *ids, int m, int r0,
                                   const float4 (&v)[G][CPL], uint64_t *keys_out) {
    constexpr int LPR = row_lanes<CPL>(), RPG = WAVE / LPR;
    const int lane = lane_id(), sub = lane % LPR, grp = lane / LPR;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        if (r0 + RPG * g >= m) break;
        const int row = r0 + RPG * g + grp;
        float acc;
        if constexpr (LPR == 16) {
            acc = 0.f;
#pragma unroll
            for (int c = 0; c < CPL; ++c)
                if (sub + 16 * c < (int)nch) acc = acc + chunk_sq(qv[c], v[g][c]);
            acc = fold16<DPP>(acc);
        } else {
            float a0 = 0.f, a1 = 0.f;  // accumulators sub and sub + 8
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                if (sub + 8 * c < (int)nch) {
                    if (c & 1) a1 = a1 + chunk_sq(qv[c], v[g][c]);
                    else a0 = a0 + chunk_sq(qv[c], v[g][c]);
                }
            }
            acc = a0 + a1;
            acc = acc + (DPP ? xor_lane(acc, 4) : __shfl_xor(acc, 4));
            acc = acc + (DPP ? xor_lane(acc, 2) : __shfl_xor(acc, 2));
            acc = acc + (DPP ? xor_lane(acc, 1) : __shfl_xor(acc, 1));
        }
        if (sub == 0 && row < m) keys_out[row] = make_key(acc, ids[row]);
    }
}

template <int CPL, int G, bool DPP = false>
__device__ inline void rows_finish(const float *__restrict__ base, uint32_t dp, uint32_t nch,
                                   const float4 (&qv)[CPL], const uint32_t *ids, int m, const float4 (&v0)[G][CPL],
                                   uint64_t *keys_out) {
    constexpr int RPG = WAVE / row_lanes<CPL>();
    rows_reduce<CPL, G, DPP>(nch, qv, ids, m, 0, v0, keys_out);
    for (int r0 = RPG * G; r0 < m; r0 += RPG * G) {
        float4 v[G][CPL];
        rows_issue<CPL, G>(base, dp, nch, ids, m, r0, v);
        rows_reduce<CPL, G, DPP>(nch, qv, ids, m, r0, v, keys_out);
    }
}

// keys_out[r] = (dist(query, base[ids[r]]), ids[r]) for r < m.  row_lanes
// lanes per row; G row groups of 64 / row_lanes rows are gathered before
// any is reduced, so that all their loads are in flight.  DPP: the folds
// exchange by DPP (see wave_sort64); the sums are the same.
template <int CPL, int G, bool DPP = false>
__device__ inline void rows_dist(const float *__restrict__ base, uint32_t dp, uint32_t nch,
                                 const float4 (&qv)[CPL], const uint32_t *ids, int m,
                                 uint64_t *keys_out) {
    constexpr int RPG = WAVE / row_lanes<CPL>();
    for (int r0 = 0; r0 < m; r0 += RPG * G) {
        float4 v[G][CPL];
        rows_issue<CPL, G>(base, dp, nch, ids, m, r0, v);
        rows_reduce<CPL, G, DPP>(nch, qv, ids, m, r0, v, keys_out);
    }
}

}  // namespace wv
