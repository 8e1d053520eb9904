// wv_nnd.hip -- NN-Descent graph construction on one GPU.
//
// Replaces ComponentInitRandom::InitInner (src/component_init.cpp:11-45),
// ComponentRefineNNDescent::RefineInner (src/component_refine.cpp:12-239)
// and ComponentPruneNaive::PruneInner (src/component_prune.cpp:8-18).
//
// State per node n, all in HBM:
//   pool[n][0..L)   ascending (dist,id) keys of its neighbour pool, unused
//                   slots KEY_EMPTY; pflag = "new" flags; pcnt = size
//   M[n]            sampling depth of update() (component_refine.cpp:167-175)
//   rec[n]          what a join reads of a listed node, in one 16-byte
//                   record: thr, the worst key of the pool (acceptance
//                   threshold), the join row's squared norm and its FP8 scale
//   worst[n]        its distance (reverse-sample test, :191/:202)
//   fnew/fold       forward new/old samples (keys) of the last update
//   rv_*            reverse proposals in CSR form, rsel_* the kept ones
//   phash[n]        open-addressing table of the pool's ids, probed by the
//                   join so that an id already pooled never takes a slot
//                   (the duplicate test of index.h:123-125, done up front)
//   cand[n][0..C)   candidate slots: a join result (id, d) for target t goes
//                   to slot hash(t, id) via 64-bit atomicMin on its key, so
//                   each slot keeps the best proposal that hashed there and
//                   the outcome does not depend on arrival order.
// Iteration it (0-based; it = 0's join is empty, component_refine.cpp:90
// with include/weavess/index.h:98-103):
//   join   one wavefront per node: new x new (i < j) and new x old Gram
//          blocks of 16-row tiles on MFMA, operands loaded straight into
//          registers, d = |a|^2 + |b|^2 - 2 a.b with precomputed squared
//          norms; proposals at or below the target's threshold queue in
//          LDS and go to its candidate slots 64 at a time;
//   merge  one wavefront per node: pool U slot winners (duplicates by id
//          dropped, pool entry wins), bitonic sort, first L kept, winners
//          flagged new, slots cleared, new M;
//   sample forward samples + reverse counts, exclusive scans, reverse
//          scatter, reverse selection (R smallest hash keys, sorted by id).
// The last iteration's sampling is skipped: it changes only flags and M,
// which nothing reads afterwards.  finalize re-ranks the first
// rerank_mult*K pool entries with exact FP32 build-order distances and
// keeps the first K (PruneNaive).
#include "wv_device.hpp"
#include "wv_nnd.hpp"

#include <algorithm>
#include <cstdio>
#include <vector>

namespace wv {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int DT_F32 = 0, DT_BF16 = 1, DT_FP8 = 2;
// Statistics counters are spread over slots picked by workgroup and summed on
// the host: a million same-address atomics per launch serialise.
constexpr uint32_t CTR_SLOTS = 1024;
__device__ inline unsigned long long *ctr_slot(unsigned long long *ctr) {
    return ctr + (size_t)(blockIdx.x & (CTR_SLOTS - 1)) * 4;
}
constexpr float FP8_MAX = 448.0f;  // largest finite e4m3 magnitude

// A node's join-time scalars, fetched with one 16-byte load per list entry
// (separate arrays cost a cache line each per random read).
struct alignas(16) NodeRec {
    uint64_t thr;  // worst key of the pool
    float norm;    // squared norm of the join row
    float scale;   // FP8 row scale (1 for FP32 / BF16)
};

struct NndDev {
    const float *base;     // N x dp FP32 rows (zero padded)
    uint32_t dp;
    const uint8_t *xq;     // join rows in join precision
    uint32_t xq_stride;    // bytes per join row
    uint32_t kdim;         // elements per join row (FP32: dp; BF16/FP8: multiple of 32)
    uint32_t N, L, S, R, K, C, rerank;
    uint32_t iter;
    uint64_t seed;
    uint64_t *pool;
    uint8_t *pflag;
    uint32_t *pcnt, *M;
    NodeRec *rec;
    float *worst;
    uint64_t *fnew, *fold;
    uint32_t *fnew_cnt, *fold_cnt;
    uint32_t *rcnt_new, *rcnt_old, *roff_new, *roff_old, *rv_new, *rv_old;
    uint32_t *rpos_new, *rpos_old;  // per fnew / fold entry: its slot in the neighbour's reverse list, or ID_NONE
    uint32_t *rsel_new, *rsel_old, *rsel_new_cnt, *rsel_old_cnt;
    uint64_t *cand;
    uint32_t *phash;          // N x HP open-addressing table of the pool's ids
    uint32_t HP;
    unsigned long long *ctr;  // CTR_SLOTS x [0] pairs [1] proposals [2] rows gathered [3] pool updates
    uint32_t *out_ids;
    float *out_dist;
    uint32_t *out_deg;
    uint32_t lds_join_list;   // bytes of list area in the join's per-wave LDS
    uint32_t nw_pow2;         // power of two >= the longest new list (S + R; 2S in EFANNA's first round)
    uint32_t fst;             // ids per fnew row: S, or 2S for EFANNA's first-round samples
    uint32_t ncap;            // longest new + old list: S + 3R, or at least 2S for EFANNA
    const uint32_t *order;    // processing order of the nodes (null: by id)
};

// The node the g-th wavefront serves in the kernels that gather other
// nodes' data (init, join, finalize): the g-th of the processing order, or
// g.  The order groups nodes by their nearest pivot row, so the wavefronts
// resident at one time work on one neighbourhood and its rows, records,
// pool-id tables and candidate slots stay in L2.  Their outcome is the same
// whichever wavefront serves a node and when: the join's slot updates are
// atomicMin, every other write is the node's own.  The kernels that work on
// a node's own rows (sample, reverse scatter and selection, merge, hash)
// keep id order, in which neighbouring wavefronts read neighbouring rows.
__device__ inline uint32_t node_at(const NndDev &a, uint32_t g) { return a.order ? a.order[g] : g; }

__device__ inline uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o);
    return v;
}

// ------------------------------------------------------------------ fill
__global__ void fill_u64_kernel(uint64_t *p, uint64_t v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void fill_u32_kernel(uint32_t *p, uint32_t v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// ------------------------------------------------------------------ scan
// Exclusive scan of n u32 counts into out[0..n] (out[n] = total): per-block
// sums, one block scanning the sums, per-block rescan with the offset.
constexpr int SCAN_T = 256, SCAN_IT = 16, SCAN_B = SCAN_T * SCAN_IT;

__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t *tmp, uint32_t &total) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = (uint32_t)__shfl_up((int)x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) tmp[w] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
        if (i < w) off += tmp[i];
        tot += tmp[i];
    }
    __syncthreads();
    total = tot;
    return off + x - v;
}

__global__ void scan_sums_kernel(const uint32_t *in, uint32_t n, uint32_t *bsum) {
    __shared__ uint32_t tmp[SCAN_T / 64];
    const size_t b0 = (size_t)blockIdx.x * SCAN_B;
    uint32_t s = 0;
    for (int i = 0; i < SCAN_IT; ++i) {
        size_t j = b0 + (size_t)i * SCAN_T + threadIdx.x;
        if (j < n) s += in[j];
    }
    uint32_t tot;
    (void)block_excl_scan(s, tmp, tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}
__global__ void scan_bsum_kernel(uint32_t *bsum, uint32_t nb) {
    __shared__ uint32_t tmp[SCAN_T / 64];
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < nb; c0 += SCAN_T) {
        uint32_t j = c0 + threadIdx.x;
        uint32_t v = j < nb ? bsum[j] : 0;
        uint32_t tot;
        uint32_t ex = block_excl_scan(v, tmp, tot);
        if (j < nb) bsum[j] = carry + ex;
        carry += tot;
    }
}
__global__ void scan_final_kernel(const uint32_t *in, uint32_t n, const uint32_t *bsum, uint32_t *out) {
    __shared__ uint32_t tmp[SCAN_T / 64];
    const size_t b0 = (size_t)blockIdx.x * SCAN_B;
    // each thread owns SCAN_IT consecutive elements
    const size_t j0 = b0 + (size_t)threadIdx.x * SCAN_IT;
    uint32_t v[SCAN_IT], s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_IT; ++i) {
        v[i] = (j0 + i < n) ? in[j0 + i] : 0;
        s += v[i];
    }
    uint32_t tot;
    uint32_t run = bsum[blockIdx.x] + block_excl_scan(s, tmp, tot);
#pragma unroll
    for (int i = 0; i < SCAN_IT; ++i) {
        if (j0 + i < n) out[j0 + i] = run;
        run += v[i];
    }
    if (j0 < n && j0 + SCAN_IT >= n) out[n] = run;  // thread holding the last element
}

// ------------------------------------------------------------------ prep
// Squared norms of the join rows; for BF16/FP8 also the quantised rows.
// 16 lanes per row, chunk c of lane j = j + 16c.
template <int DT>
__global__ void nnd_prep_kernel(NndDev a, uint8_t *xq_out) {
    const uint32_t row = (blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    const int sub = threadIdx.x & 15;
    const bool live = row < a.N;
    const uint32_t r = live ? row : 0;
    const float4 *src = reinterpret_cast<const float4 *>(a.base + (size_t)r * a.dp);
    const uint32_t nsrc = a.dp / 4;           // chunks present in the FP32 row
    const uint32_t nch = (DT == DT_F32) ? nsrc : a.kdim / 4;
    float acc = 0.f, amax = 0.f;
    if (DT == DT_FP8) {
        for (uint32_t c = sub; c < nsrc; c += 16) {
            float4 v = src[c];
            amax = fmaxf(amax, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
        }
        amax = fmaxf(amax, __shfl_xor(amax, 8));
        amax = fmaxf(amax, __shfl_xor(amax, 4));
        amax = fmaxf(amax, __shfl_xor(amax, 2));
        amax = fmaxf(amax, __shfl_xor(amax, 1));
    }
    const float sc = (DT == DT_FP8) ? (amax > 0.f ? amax / FP8_MAX : 1.0f) : 1.0f;
    const float inv = 1.0f / sc;
    for (uint32_t c = sub; c < nch; c += 16) {
        float4 v = c < nsrc ? src[c] : make_float4(0.f, 0.f, 0.f, 0.f);
        if (DT == DT_F32) {
            acc = acc + (v.x * v.x + v.y * v.y + v.z * v.z + v.w * v.w);
        } else if (DT == DT_BF16) {
            __bf16 b0 = (__bf16)v.x, b1 = (__bf16)v.y, b2 = (__bf16)v.z, b3 = (__bf16)v.w;
            float f0 = (float)b0, f1 = (float)b1, f2 = (float)b2, f3 = (float)b3;
            acc = acc + (f0 * f0 + f1 * f1 + f2 * f2 + f3 * f3);
            if (live) {
                __bf16 *dst = reinterpret_cast<__bf16 *>(xq_out + (size_t)r * a.xq_stride) + 4 * c;
                dst[0] = b0; dst[1] = b1; dst[2] = b2; dst[3] = b3;
            }
        } else {
            float x0 = fminf(fmaxf(v.x * inv, -FP8_MAX), FP8_MAX), x1 = fminf(fmaxf(v.y * inv, -FP8_MAX), FP8_MAX);
            float x2 = fminf(fmaxf(v.z * inv, -FP8_MAX), FP8_MAX), x3 = fminf(fmaxf(v.w * inv, -FP8_MAX), FP8_MAX);
            int w = __builtin_amdgcn_cvt_pk_fp8_f32(x0, x1, 0, false);
            w = __builtin_amdgcn_cvt_pk_fp8_f32(x2, x3, w, true);
            float f0 = __builtin_amdgcn_cvt_f32_fp8(w, 0), f1 = __builtin_amdgcn_cvt_f32_fp8(w, 1);
            float f2 = __builtin_amdgcn_cvt_f32_fp8(w, 2), f3 = __builtin_amdgcn_cvt_f32_fp8(w, 3);
            acc = acc + (f0 * f0 + f1 * f1 + f2 * f2 + f3 * f3);
            if (live) *reinterpret_cast<int *>(xq_out + (size_t)r * a.xq_stride + 4 * c) = w;
        }
    }
    acc = fold16(acc);
    if (live && sub == 0) {
        a.rec[r].norm = (DT == DT_FP8) ? acc * sc * sc : acc;
        a.rec[r].scale = sc;
    }
}

// ------------------------------------------------------------------ M
// update()'s sampling depth, src/component_refine.cpp:167-175: scan from the
// front while l < min(M+S, size) and fewer than S new entries were seen.
__device__ inline uint32_t compute_M(const uint8_t *flags, uint32_t Mprev, uint32_t S, uint32_t cnt) {
    const int lane = lane_id();
    const uint32_t maxl = min(Mprev + S, cnt);
    uint32_t c = 0;
    if (S == 0) return 0;
    for (uint32_t t0 = 0; t0 < maxl; t0 += WAVE) {
        const uint32_t p = t0 + lane;
        const uint64_t b = __ballot(p < maxl && flags[p]);
        const uint32_t pc = (uint32_t)__popcll(b);
        if (c + pc >= S) {
            uint64_t m = b;
            for (uint32_t i = 1; i < S - c; ++i) m &= m - 1;
            return t0 + (uint32_t)__ffsll((unsigned long long)m);  // position + 1
        }
        c += pc;
    }
    return maxl;
}

// ------------------------------------------------------------------ init
// GenRandom (include/weavess/util.h:13-27) of size <= 64 on counter stream
// s: lane i < size returns the i-th id (ascending before the rotation),
// other lanes ID_NONE.  The draws are reduced mod (N - size), sorted and
// forced strictly increasing by a prefix max of (v_i - i), then rotated by
// one more draw mod N.
__device__ inline uint32_t wave_gen_random(uint64_t s, uint32_t size, uint32_t N) {
    const int lane = lane_id();
    uint64_t v = (uint32_t)lane < size ? (uint64_t)(draw(s, lane) % (N - size)) : KEY_EMPTY;
    v = wave_sort64(v);
    const int64_t b = (uint32_t)lane < size ? (int64_t)v - lane : INT64_MIN;
    const int64_t pm = wave_scan_max64(b);
    const uint64_t off = draw(s, size) % N;
    return (uint32_t)lane < size ? (uint32_t)(((uint64_t)(pm + lane) + off) % N) : ID_NONE;
}

// keys_out[r] = (d, ids[r]) for r < m with d node n's distance to ids[r]
// by the join's formula on the join-precision rows: |a|^2 + |b|^2 - 2 a.b
// (the FP8 dot times both rows' scales), FP32 products and sums, clamped at
// 0.  Four lanes a row, 16 rows a pass; lane s of a row takes the row's
// 16-byte vectors s, s + 4, ...; rows past m read node n's own row and are
// not stored.
template <int DT>
__device__ inline void lp_rows_dist(const NndDev &a, uint32_t n, const uint32_t *ids, int m, uint64_t *keys_out) {
    const int lane = lane_id(), sub = lane & 3, grp = lane >> 2;
    const uint32_t nv = a.xq_stride / 16;  // whole vectors: kdim is a multiple of 32
    const uint4 *qa = reinterpret_cast<const uint4 *>(a.xq + (size_t)n * a.xq_stride);
    const float na = a.rec[n].norm, sa = a.rec[n].scale;
    for (int r0 = 0; r0 < m; r0 += 16) {
        const int row = r0 + grp;
        const bool live = row < m;
        const uint32_t id = live ? ids[row] : n;
        const uint4 *qb = reinterpret_cast<const uint4 *>(a.xq + (size_t)id * a.xq_stride);
        float acc = 0.f;
#pragma unroll 4
        for (uint32_t v = (uint32_t)sub; v < nv; v += 4) {
            const uint4 x = qa[v], y = qb[v];
            const uint32_t xs[4] = {x.x, x.y, x.z, x.w}, ys[4] = {y.x, y.y, y.z, y.w};
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                if (DT == DT_BF16) {
                    acc = acc + __uint_as_float(xs[w] << 16) * __uint_as_float(ys[w] << 16);
                    acc = acc + __uint_as_float(xs[w] & 0xFFFF0000u) * __uint_as_float(ys[w] & 0xFFFF0000u);
                } else {
                    const int xw = (int)xs[w], yw = (int)ys[w];
                    acc = acc + __builtin_amdgcn_cvt_f32_fp8(xw, 0) * __builtin_amdgcn_cvt_f32_fp8(yw, 0);
                    acc = acc + __builtin_amdgcn_cvt_f32_fp8(xw, 1) * __builtin_amdgcn_cvt_f32_fp8(yw, 1);
                    acc = acc + __builtin_amdgcn_cvt_f32_fp8(xw, 2) * __builtin_amdgcn_cvt_f32_fp8(yw, 2);
                    acc = acc + __builtin_amdgcn_cvt_f32_fp8(xw, 3) * __builtin_amdgcn_cvt_f32_fp8(yw, 3);
                }
            }
        }
        acc = acc + __shfl_xor(acc, 1);
        acc = acc + __shfl_xor(acc, 2);
        if (sub == 0 && live) {
            float dot = acc;
            if (DT == DT_FP8) dot = dot * (sa * a.rec[id].scale);
            const float d = fmaxf((na + a.rec[id].norm) - 2.0f * dot, 0.0f);
            keys_out[row] = make_key(d, id);
        }
    }
}

// InitRandom: per node, S draws of GenRandom on the node's stream (as in
// SearchEntryRand, a prefix-max scan forces them strictly increasing), self
// skipped, distances, ascending pool, all new, M = S.  FP32 builds take
// build-order FP32 distances (rows of CPL chunks a lane); BF16 and FP8
// builds the join's formula on their join rows (half or a quarter of the
// bytes), as every later distance in their pools.
template <int CPL, int DT>
__global__ void __launch_bounds__(256) nnd_init_kernel(NndDev a, const uint32_t *init_ids,
                                                       const uint32_t *init_cnt) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int wid = threadIdx.x >> 6, lane = lane_id();
    const uint32_t g = blockIdx.x * (blockDim.x >> 6) + wid;
    if (g >= a.N) return;
    const uint32_t n = node_at(a, g);
    uint32_t *cid = reinterpret_cast<uint32_t *>(smem) + wid * 64;
    uint64_t *ckey = reinterpret_cast<uint64_t *>(smem + (blockDim.x >> 6) * 256) + wid * 64;
    const uint32_t S = a.S, N = a.N;
    uint32_t id = ID_NONE;
    if (init_ids) {
        const uint32_t c = init_cnt ? init_cnt[n] : S;
        if ((uint32_t)lane < c && (uint32_t)lane < S) id = init_ids[(size_t)n * S + lane];
    } else {
        id = wave_gen_random(stream_of(a.seed, TAG_INIT, n), S, N);
    }
    const bool ok = id < N && id != n;
    const uint64_t bm = __ballot(ok);
    const int m = __popcll(bm);
    if (ok) cid[lanes_below(bm)] = id;
    wave_sync();
    if constexpr (DT == DT_F32) {
        float4 qv[CPL];
        const float4 *qp = reinterpret_cast<const float4 *>(a.base + (size_t)n * a.dp);
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            const int ch = (lane & 15) + 16 * c;
            qv[c] = ch < (int)(a.dp / 4) ? qp[ch] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        // long rows gather one row group at a time: four groups of 8-16
        // chunks a lane would take 180-256 VGPRs and leave one or two
        // wavefronts a SIMD (the sums are the same either way)
        rows_dist<CPL, (CPL >= 8 ? 1 : 4)>(a.base, a.dp, a.dp / 4, qv, cid, m, ckey);
    } else {
        lp_rows_dist<DT>(a, n, cid, m, ckey);
    }
    wave_sync();
    uint64_t key = lane < m ? ckey[lane] : KEY_EMPTY;
    key = wave_sort64(key);
    const uint64_t last = shfl64(key, m > 0 ? m - 1 : 0);
    // slot i < 64 is lane i's sorted key (KEY_EMPTY past m); the rest empty
    for (uint32_t i = lane; i < a.L; i += WAVE) {
        a.pool[(size_t)n * a.L + i] = i < 64 ? key : KEY_EMPTY;
        a.pflag[(size_t)n * a.L + i] = i < (uint32_t)m ? 1 : 0;
    }
    if (lane == 0) {
        a.pcnt[n] = (uint32_t)m;
        a.M[n] = S;
        a.rec[n].thr = m > 0 ? last : KEY_EMPTY;
        a.worst[n] = m > 0 ? key_dist(last) : -1.0f;
    }
}

__global__ void __launch_bounds__(256) nnd_m0_kernel(NndDev a) {
    const int wid = threadIdx.x >> 6;
    const uint32_t n = blockIdx.x * (blockDim.x >> 6) + wid;
    if (n >= a.N) return;
    const uint32_t m = compute_M(a.pflag + (size_t)n * a.L, a.M[n], a.S, a.pcnt[n]);
    if (lane_id() == 0) a.M[n] = m;
}

// Per-round monitor (ComponentRefineNNDescent::eval_recall,
// src/component_refine.cpp:258-275): one wavefront per control node counts
// the pool's ids that appear in the node's truth row (held in LDS).
__global__ void __launch_bounds__(256) nnd_monitor_kernel(NndDev a, NndMonitor m, uint32_t round) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int wid = threadIdx.x >> 6, lane = lane_id();
    const uint32_t i = blockIdx.x * (blockDim.x >> 6) + wid;
    if (i >= m.n_ctrl) return;
    uint32_t *tr = reinterpret_cast<uint32_t *>(smem) + (size_t)wid * m.truth_k;
    for (uint32_t j = lane; j < m.truth_k; j += WAVE) tr[j] = m.truth[(size_t)i * m.truth_k + j];
    wave_sync();
    const uint32_t c = m.ctrl[i];
    const uint64_t *pr = a.pool + (size_t)c * a.L;
    uint32_t hit = 0;
    for (uint32_t j = lane; j < a.L; j += WAVE) {
        const uint64_t key = pr[j];
        if (key == KEY_EMPTY) continue;
        const uint32_t id = key_id(key);
        for (uint32_t t = 0; t < m.truth_k; ++t)
            if (tr[t] == id) {
                ++hit;
                break;
            }
    }
    hit = wave_sum_u32(hit);
    if (lane == 0) m.hits[(size_t)round * m.n_ctrl + i] = hit;
}

// ComponentRefineEFANNA::init (src/component_refine.cpp:1036-1058): the pool
// is the initial graph's row (up to `stride` <= L ids, cnt[n] valid) sorted
// by (dist, id) with self, out-of-range and repeated ids dropped, first L
// kept, all new, M = S.  The row's keys are gathered 64 at a time into a
// P-key LDS array (P = power of two >= stride), sorted, repeats (equal keys,
// hence adjacent) cleared and sorted again.
template <int CPL>
__global__ void __launch_bounds__(256) nnd_init_rows_kernel(NndDev a, const uint32_t *ids, const uint32_t *cnt,
                                                            uint32_t stride, uint32_t P) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int wid = threadIdx.x >> 6, lane = lane_id();
    const int wpb = blockDim.x >> 6;
    const uint32_t n = blockIdx.x * wpb + wid;
    if (n >= a.N) return;
    uint32_t *cid = reinterpret_cast<uint32_t *>(smem) + wid * 64;
    uint64_t *ckey = reinterpret_cast<uint64_t *>(smem + wpb * 256) + wid * 64;
    uint64_t *keys = reinterpret_cast<uint64_t *>(smem + wpb * (256 + 512)) + (size_t)wid * P;
    const uint32_t N = a.N;
    const uint32_t c = cnt ? min(cnt[n], stride) : stride;
    float4 qv[CPL];
    {
        const float4 *qp = reinterpret_cast<const float4 *>(a.base + (size_t)n * a.dp);
#pragma unroll
        for (int k = 0; k < CPL; ++k) {
            const int ch = (lane & 15) + 16 * k;
            qv[k] = ch < (int)(a.dp / 4) ? qp[ch] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
    uint32_t tot = 0;
    for (uint32_t c0 = 0; c0 < c; c0 += WAVE) {
        const uint32_t id = c0 + lane < c ? ids[(size_t)n * stride + c0 + lane] : ID_NONE;
        const bool ok = id < N && id != n;
        const uint64_t bm = __ballot(ok);
        const int m = __popcll(bm);
        if (ok) cid[lanes_below(bm)] = id;
        wave_sync();
        rows_dist<CPL, (CPL >= 8 ? 1 : 4)>(a.base, a.dp, a.dp / 4, qv, cid, m, ckey);  // as in nnd_init_kernel
        wave_sync();
        if (lane < m) keys[tot + lane] = ckey[lane];
        tot += (uint32_t)m;
        wave_sync();
    }
    for (uint32_t i = tot + lane; i < P; i += WAVE) keys[i] = KEY_EMPTY;
    wave_sync();
    lds_sort64(keys, (int)P);
    uint32_t rep = 0;  // bit j: slot lane + 64 j repeats its predecessor (P <= 1024)
    for (uint32_t j = 0; j * WAVE < P; ++j) {
        const uint32_t i = lane + j * WAVE;
        if (i > 0 && keys[i] != KEY_EMPTY && keys[i] == keys[i - 1]) rep |= 1u << j;
    }
    wave_sync();
    for (uint32_t j = 0; j * WAVE < P; ++j)
        if (rep >> j & 1u) keys[lane + j * WAVE] = KEY_EMPTY;
    wave_sync();
    if (__ballot(rep != 0)) lds_sort64(keys, (int)P);
    uint32_t uniq = 0;
    for (uint32_t j = 0; j * WAVE < P; ++j) uniq += (uint32_t)__popcll(__ballot(keys[lane + j * WAVE] != KEY_EMPTY));
    const uint32_t m = min(uniq, a.L);
    for (uint32_t i = lane; i < a.L; i += WAVE) {
        a.pool[(size_t)n * a.L + i] = i < m ? keys[i] : KEY_EMPTY;
        a.pflag[(size_t)n * a.L + i] = i < m ? 1 : 0;
    }
    if (lane == 0) {
        const uint64_t last = m > 0 ? keys[m - 1] : KEY_EMPTY;
        a.pcnt[n] = m;
        a.M[n] = a.S;
        a.rec[n].thr = last;
        a.worst[n] = m > 0 ? key_dist(last) : -1.0f;
    }
}

// EFANNA's first-round lists (include/weavess/index.h:105-110): nn_new is
// GenRandom(2S) over [0, N) on the node's "EFNA" stream (it may hold the
// node itself, as in the reference), nn_old and the reverse lists empty.
// 2S <= 128: the draws are sorted in a 128-key LDS array, and lane l forces
// elements 2l and 2l+1 strictly increasing with the prefix max of (v_i - i)
// carried in from the lanes below.
__global__ void __launch_bounds__(256) nnd_efanna_seed_kernel(NndDev a) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int wid = threadIdx.x >> 6, lane = lane_id();
    const uint32_t n = blockIdx.x * (blockDim.x >> 6) + wid;
    if (n >= a.N) return;
    uint64_t *v = reinterpret_cast<uint64_t *>(smem) + wid * 128;
    const uint32_t S2 = 2 * a.S, N = a.N;
    const uint64_t st = stream_of(a.seed, TAG_EFANNA, n);
    for (uint32_t i = lane; i < 128; i += WAVE) v[i] = i < S2 ? (uint64_t)(draw(st, i) % (N - S2)) : KEY_EMPTY;
    wave_sync();
    lds_sort64(v, 128);
    const uint32_t i0 = 2 * lane, i1 = i0 + 1;
    int64_t mx = INT64_MIN;
    if (i0 < S2) mx = (int64_t)v[i0] - (int64_t)i0;
    if (i1 < S2) mx = max(mx, (int64_t)v[i1] - (int64_t)i1);
    const int64_t incl = wave_scan_max64(mx);
    const uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)incl, 1);
    const uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)((uint64_t)incl >> 32), 1);
    int64_t run = lane == 0 ? INT64_MIN : (int64_t)(((uint64_t)hi << 32) | lo);
    const uint64_t off = draw(st, S2) % N;
    uint64_t *dst = a.fnew + (size_t)n * a.fst;
    if (i0 < S2) {
        run = max(run, (int64_t)v[i0] - (int64_t)i0);
        dst[i0] = ((uint64_t)(run + i0) + off) % N;
    }
    if (i1 < S2) {
        run = max(run, (int64_t)v[i1] - (int64_t)i1);
        dst[i1] = ((uint64_t)(run + i1) + off) % N;
    }
    if (lane == 0) {
        a.fnew_cnt[n] = S2;
        a.fold_cnt[n] = 0;
        a.rsel_new_cnt[n] = 0;
        a.rsel_old_cnt[n] = 0;
    }
}

// ------------------------------------------------------------------ sample
// update() :176-209 forward part: first M pool entries, flagged ones to the
// new list (flag cleared), the rest to the old list, in pool order; count a
// reverse proposal at the neighbour when dist > neighbour's worst, and keep
// the count's old value as the proposal's slot in the neighbour's list.
__global__ void __launch_bounds__(256) nnd_sample_kernel(NndDev a) {
    const int wid = threadIdx.x >> 6, lane = lane_id();
    const uint32_t n = blockIdx.x * (blockDim.x >> 6) + wid;
    if (n >= a.N) return;
    const uint32_t Mn = a.M[n];
    const uint64_t *pr = a.pool + (size_t)n * a.L;
    uint8_t *fl = a.pflag + (size_t)n * a.L;
    // the first 64 slots load with M, not after it (a row holds L slots;
    // those at or past M are masked below)
    const uint64_t key0 = (uint32_t)lane < a.L ? pr[lane] : KEY_EMPTY;
    const uint8_t f0 = (uint32_t)lane < a.L ? fl[lane] : (uint8_t)0;
    uint32_t nf = 0, no = 0;
    for (uint32_t t0 = 0; t0 < Mn; t0 += WAVE) {
        const uint32_t p = t0 + lane;
        const bool act = p < Mn;
        const uint64_t key = act ? (t0 == 0 ? key0 : pr[p]) : KEY_EMPTY;
        const bool f = act && (t0 == 0 ? f0 : fl[p]);
        const uint64_t bn = __ballot(f), bo = __ballot(act && !f);
        if (act) {
            const uint32_t id = key_id(key);
            const uint32_t pos =
                key_dist(key) > a.worst[id] ? atomicAdd(f ? &a.rcnt_new[id] : &a.rcnt_old[id], 1u) : ID_NONE;
            if (f) {
                const size_t j = (size_t)n * a.fst + nf + lanes_below(bn);
                a.fnew[j] = key;
                a.rpos_new[j] = pos;
            } else {
                const size_t j = (size_t)n * a.L + no + lanes_below(bo);
                a.fold[j] = key;
                a.rpos_old[j] = pos;
            }
            fl[p] = 0;
        }
        nf += (uint32_t)__popcll(bn);
        no += (uint32_t)__popcll(bo);
    }
    if (lane == 0) {
        a.fnew_cnt[n] = nf;
        a.fold_cnt[n] = no;
    }
}

// Place each reverse proposal at its slot in the neighbour's list, the
// neighbour's offset plus the slot the sample kernel's count handed out.
__global__ void __launch_bounds__(256) nnd_rev_scatter_kernel(NndDev a) {
    const int wid = threadIdx.x >> 6, lane = lane_id();
    const uint32_t n = blockIdx.x * (blockDim.x >> 6) + wid;
    if (n >= a.N) return;
    const uint32_t nf = a.fnew_cnt[n], no = a.fold_cnt[n];
    // the first 64 entries of both lists load with the counts, not after
    // them (rows hold fst and L entries; those past a count are masked)
    const size_t bn = (size_t)n * a.fst, bo = (size_t)n * a.L;
    const uint32_t pn0 = (uint32_t)lane < a.fst ? a.rpos_new[bn + lane] : ID_NONE;
    const uint64_t kn0 = (uint32_t)lane < a.fst ? a.fnew[bn + lane] : 0;
    const uint32_t po0 = (uint32_t)lane < a.L ? a.rpos_old[bo + lane] : ID_NONE;
    const uint64_t ko0 = (uint32_t)lane < a.L ? a.fold[bo + lane] : 0;
    for (uint32_t i = lane; i < nf; i += WAVE) {
        const uint32_t pos = i < (uint32_t)WAVE ? pn0 : a.rpos_new[bn + i];
        if (pos == ID_NONE) continue;
        const uint32_t id = key_id(i < (uint32_t)WAVE ? kn0 : a.fnew[bn + i]);
        a.rv_new[a.roff_new[id] + pos] = n;
    }
    for (uint32_t i = lane; i < no; i += WAVE) {
        const uint32_t pos = i < (uint32_t)WAVE ? po0 : a.rpos_old[bo + i];
        if (pos == ID_NONE) continue;
        const uint32_t id = key_id(i < (uint32_t)WAVE ? ko0 : a.fold[bo + i]);
        a.rv_old[a.roff_old[id] + pos] = n;
    }
}

// Reverse-sample priority key of proposer p at node x (hash high word, p low
// word); the R smallest are kept.
__device__ inline uint64_t rev_stream(uint64_t seed, uint64_t tag, uint32_t iter, uint32_t x) {
    return stream_of(seed, tag, ((uint64_t)iter << 32) | x);
}
__device__ inline uint64_t rev_key(uint64_t sx, uint32_t p) {
    uint64_t h = fmix64(sx ^ ((uint64_t)p * GOLDEN));
    return ((h >> 32) << 32) | p;
}

// Keep at most R reverse proposals of node x (update() :190-208, :213-228):
// all when there are at most R, else the R smallest rev_key (radix select
// of the R-th smallest over the segment, a byte per pass, stopping once every
// key of the chosen bin is needed); the kept ids are sorted ascending.
__global__ void __launch_bounds__(256) nnd_rev_select_kernel(NndDev a, uint32_t rp2) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int wid = threadIdx.x >> 6, lane = lane_id();
    const uint32_t x = blockIdx.x * (blockDim.x >> 6) + wid;
    if (x >= a.N) return;
    uint8_t *my = smem + (size_t)wid * (rp2 * 4 + 256 * 4 + 16);
    uint32_t *ids = reinterpret_cast<uint32_t *>(my);
    uint32_t *hist = ids + rp2;
    uint32_t *cnt = hist + 256;
    // both lists' offsets, then the first 64 ids of both segments, before
    // either is processed
    const uint32_t on0 = a.roff_new[x], on1 = a.roff_new[x + 1];
    const uint32_t oo0 = a.roff_old[x], oo1 = a.roff_old[x + 1];
    const uint32_t vfirst[2] = {(uint32_t)lane < on1 - on0 ? a.rv_new[on0 + lane] : ID_NONE,
                                (uint32_t)lane < oo1 - oo0 ? a.rv_old[oo0 + lane] : ID_NONE};
    for (int which = 0; which < 2; ++which) {
        const uint32_t *seg = which ? a.rv_old + oo0 : a.rv_new + on0;
        const uint32_t c = which ? oo1 - oo0 : on1 - on0;
        const uint64_t sx = rev_stream(a.seed, which ? TAG_REVOLD : TAG_REVNEW, a.iter, x);
        const uint32_t R = a.R;
        uint32_t kept;
        if (c <= R && c <= (uint32_t)WAVE) {
            // all kept, one per lane: sorted in registers
            int w = 2;
            while ((uint32_t)w < c) w <<= 1;
            const uint64_t v = wave_sort64_n((uint32_t)lane < c ? (uint64_t)vfirst[which] : (uint64_t)ID_NONE, w);
            if ((uint32_t)lane < c) (which ? a.rsel_old : a.rsel_new)[(size_t)x * R + lane] = (uint32_t)v;
            if (lane == 0) (which ? a.rsel_old_cnt : a.rsel_new_cnt)[x] = c;
            continue;
        }
        if (c <= R) {
            for (uint32_t i = lane; i < rp2; i += WAVE) ids[i] = i < c ? seg[i] : ID_NONE;
            kept = c;
        } else {
            uint64_t prefix = 0, pmask = 0;
            uint32_t need = R;
            for (int shift = 56; shift >= 0; shift -= 8) {
                for (int i = lane; i < 256; i += WAVE) hist[i] = 0;
                wave_sync();
                for (uint32_t i = lane; i < c; i += WAVE) {
                    const uint64_t k = rev_key(sx, seg[i]);
                    if ((k & pmask) == prefix) atomicAdd(&hist[(k >> shift) & 255], 1u);
                }
                wave_sync();
                // bin where the running count reaches `need`
                uint32_t h4[4], s4 = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    h4[j] = hist[lane * 4 + j];
                    s4 += h4[j];
                }
                uint32_t incl = s4;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    uint32_t y = (uint32_t)__shfl_up((int)incl, o);
                    if (lane >= o) incl += y;
                }
                uint32_t before = incl - s4;
                int bin = -1;
                uint32_t below = 0;
                if (before < need && need <= incl) {
                    uint32_t run = before;
                    for (int j = 0; j < 4; ++j) {
                        if (run + h4[j] >= need) {
                            bin = lane * 4 + j;
                            below = run;
                            break;
                        }
                        run += h4[j];
                    }
                }
                const uint64_t bb = __ballot(bin >= 0);
                const int src = __ffsll((unsigned long long)bb) - 1;
                bin = __shfl(bin, src);
                below = (uint32_t)__shfl((int)below, src);
                prefix |= (uint64_t)bin << shift;
                pmask |= (uint64_t)255 << shift;
                need -= below;
                const uint32_t inbin = hist[bin];
                wave_sync();
                if (inbin == need) break;  // the whole bin is kept: no finer pass needed
            }
            // keys are distinct (low word = proposer id): exactly R have
            // (key & pmask) <= prefix
            if (lane == 0) *cnt = 0;
            for (uint32_t i = lane; i < rp2; i += WAVE) ids[i] = ID_NONE;
            wave_sync();
            for (uint32_t i = lane; i < c; i += WAVE) {
                const uint32_t p = seg[i];
                if ((rev_key(sx, p) & pmask) <= prefix) {
                    const uint32_t pos = atomicAdd(cnt, 1u);
                    if (pos < rp2) ids[pos] = p;
                }
            }
            kept = R;
        }
        wave_sync();
        int sw = 2;  // slots from `kept` up hold ID_NONE: sort the power of two covering it
        while ((uint32_t)sw < kept) sw <<= 1;
        lds_sort32(ids, sw);
        uint32_t *dst = (which ? a.rsel_old : a.rsel_new) + (size_t)x * R;
        for (uint32_t i = lane; i < kept; i += WAVE) dst[i] = ids[i];
        if (lane == 0) (which ? a.rsel_old_cnt : a.rsel_new_cnt)[x] = kept;
        wave_sync();
    }
}

// ------------------------------------------------------------------ join
// Per element type: bytes per element, elements per MFMA step, the lane's
// operand vector of one step (KSTEP / 4 elements), and the steps per k
// piece of rows longer than 512 bytes (32 bytes a lane of each of the four
// tiles a block holds in registers).
template <int DT> struct JoinT;
template <> struct JoinT<DT_F32> {
    static constexpr int ES = 4, KSTEP = 16, PIECE = 2;
    using V = uint4;
};
template <> struct JoinT<DT_BF16> {
    static constexpr int ES = 2, KSTEP = 32, PIECE = 2;
    using V = uint4;
};
template <> struct JoinT<DT_FP8> {
    static constexpr int ES = 1, KSTEP = 32, PIECE = 4;
    using V = uint2;
};

// Join rows go to the MFMA straight from HBM, without LDS: in the 16x16
// operand layout lane (r = lane & 15, g = lane >> 4) feeds row r's elements
// [KSTEP t + KSTEP/4 g, + KSTEP/4) at step t, bytes no other lane reads.
// A whole row tile of TS steps (rows of at most 512 bytes), branch-free so
// that waits on these loads can count them: rows at or past `count` read
// row 0 and steps at or past T repeat step T - 1; their products reach only
// outputs no proposal reads or steps no MFMA takes.
// FP32 elements at or past kdim (padded widths that are not a multiple of
// 16) read the row's first vector and are zeroed at the MFMA (ops_mfma_k);
// BF16 and FP8 rows are padded to a multiple of 32 and have none.
template <int DT, int TS>
__device__ inline void load_tile(const NndDev &a, const uint32_t *list, int r0, int count, int T,
                                 typename JoinT<DT>::V ops[TS]) {
    using V = typename JoinT<DT>::V;
    constexpr int KSTEP = JoinT<DT>::KSTEP, ES = JoinT<DT>::ES;
    const int lane = lane_id(), r = lane & 15, g = lane >> 4;
    const uint32_t li = list[r0 + r];  // LDS; read past the list, unused there
    const uint8_t *row = a.xq + (size_t)(r0 + r < count ? li : 0u) * a.xq_stride;
#pragma unroll
    for (int j = 0; j < TS; ++j) {
        const uint32_t e0 = (uint32_t)(KSTEP * min(j, T - 1) + (KSTEP / 4) * g);
        ops[j] = *reinterpret_cast<const V *>(row + (size_t)(e0 < a.kdim ? e0 : 0u) * ES);
    }
}

// A k piece of a tile for the streamed form: steps t0 .. t0+PS-1 of rows
// list[r0 .. r0+16), branch-free like load_tile (rows at or past `count`
// read row 0, steps at or past T repeat step T - 1 and are not multiplied,
// FP32 elements at or past kdim read the row's first vector and are zeroed
// at the MFMA).  Tiles past the list read list entries past its end, at
// most 64 (the block's last tile): LDS of the same wavefront's region,
// which holds the thresholds, norms and proposal ring after the list, and
// the values read there are never used.
template <int DT, int PS>
__device__ inline void load_piece(const NndDev &a, const uint32_t *list, int r0, int count, int T, int t0,
                                  typename JoinT<DT>::V ops[PS]) {
    using V = typename JoinT<DT>::V;
    constexpr int KSTEP = JoinT<DT>::KSTEP, ES = JoinT<DT>::ES;
    const int lane = lane_id(), r = lane & 15, g = lane >> 4;
    const uint32_t li = list[r0 + r];
    const uint8_t *row = a.xq + (size_t)(r0 + r < count ? li : 0u) * a.xq_stride;
#pragma unroll
    for (int j = 0; j < PS; ++j) {
        const uint32_t e0 = (uint32_t)(KSTEP * min(t0 + j, T - 1) + (KSTEP / 4) * g);
        ops[j] = *reinterpret_cast<const V *>(row + (size_t)(e0 < a.kdim ? e0 : 0u) * ES);
    }
}

// MFMA steps j < n of register operands into acc, in step order.
// ztail: this lane's step n - 1 is zero on both operands.
template <int DT, int TS>
__device__ inline floatx4 ops_mfma(const typename JoinT<DT>::V A[TS], const typename JoinT<DT>::V B[TS], int n,
                                   floatx4 acc, bool ztail = false) {
#pragma unroll
    for (int t = 0; t < TS; ++t) {
        if (t >= n) continue;
        if (DT == DT_F32) {
            float4 x, y;
            __builtin_memcpy(&x, &A[t], 16);
            __builtin_memcpy(&y, &B[t], 16);
            if (ztail && t == n - 1) x = y = make_float4(0.f, 0.f, 0.f, 0.f);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x.x, y.x, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x.y, y.y, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x.z, y.z, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x.w, y.w, acc, 0, 0, 0);
        } else if (DT == DT_BF16) {
            bf16x8 x, y;
            __builtin_memcpy(&x, &A[t], 16);
            __builtin_memcpy(&y, &B[t], 16);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x, y, acc, 0, 0, 0);
        } else {
            long x, y;
            __builtin_memcpy(&x, &A[t], 8);
            __builtin_memcpy(&y, &B[t], 8);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(x, y, acc, 0, 0, 0);
        }
    }
    return acc;
}

// ops_mfma for tiles from load_tile: FP32 step T - 1's elements at or past
// kdim are zeroed on both operands (they add nothing), only when the
// padded width is not a multiple of 16 (a wavefront-uniform test).
template <int DT, int TS>
__device__ inline floatx4 ops_mfma_k(const typename JoinT<DT>::V A[TS], const typename JoinT<DT>::V B[TS], int T,
                                     uint32_t kdim, floatx4 acc) {
    if constexpr (DT == DT_F32) {
        if (kdim & (JoinT<DT>::KSTEP - 1)) {
            const uint32_t e_last =
                (uint32_t)(JoinT<DT>::KSTEP * (T - 1) + (JoinT<DT>::KSTEP / 4) * (lane_id() >> 4));
            return ops_mfma<DT, TS>(A, B, T, acc, e_last >= kdim);
        }
    }
    return ops_mfma<DT, TS>(A, B, T, acc);
}

// Is `id` in target t's pool?  Linear probing of t's pool-id hash from slot
// h, read as aligned groups of 4 slots (HP is a multiple of 64, so groups
// never straddle the wrap): slots h..(h|3) of the first group, then whole
// groups, in probe order, until `id` or an empty slot.  Most lookups miss
// and end within the first group, in one load instead of a chain.
__device__ inline bool pooled(const NndDev &a, uint32_t t, uint32_t id, uint32_t h) {
    const uint32_t *tab = a.phash + (size_t)t * a.HP;
    const uint32_t mask = a.HP - 1;
    uint32_t g0 = h & ~3u, from = 0xFu << (h & 3u);
    for (;;) {
        const uint4 g = *reinterpret_cast<const uint4 *>(tab + g0);
        const uint32_t hit = (g.x == id ? 1u : 0u) | (g.y == id ? 2u : 0u) | (g.z == id ? 4u : 0u) |
                             (g.w == id ? 8u : 0u);
        const uint32_t emp = (g.x == ID_NONE ? 1u : 0u) | (g.y == ID_NONE ? 2u : 0u) |
                             (g.z == ID_NONE ? 4u : 0u) | (g.w == ID_NONE ? 8u : 0u);
        const uint32_t stop = (hit | emp) & from;
        if (stop) return (hit & (stop & (0u - stop))) != 0;  // the first stop in probe order
        g0 = (g0 + 4) & mask;
        from = 0xFu;
    }
}

// Proposals that pass the target's threshold wait in a per-wavefront LDS
// ring and are resolved 64 at a time, one per lane: the pool-id hash probe,
// then atomicMin on the candidate slot.  Only a few percent pass after the
// first rounds, so the wavefront probes full rows of proposals instead of
// eight sparse ones per tile.  An entry holds the distance's bits and the
// list positions of the target (low 16 bits) and of the proposed id.
// A flush is split in two: `issue` loads each lane's first hash group and
// `finish` (at the next flush, or at the end) probes and proposes, so the
// load's latency overlaps the tiles in between.  The ring holds the block
// in flight ([head, head + pn)) and the waiting entries behind it.
constexpr uint32_t JQ_CAP = 256;  // >= 128 in flight or waiting before a row's pushes, + 128 pushed
constexpr uint32_t JQ_BYTES = JQ_CAP * 8;
struct JoinQ {
    uint2 *q;
    uint32_t head, pn, n;  // wavefront-uniform
    uint4 g;               // this lane's first hash group of the block in flight
};

__device__ inline void jq_push(JoinQ &Q, bool v, float d, uint32_t ti, uint32_t si) {
    const uint64_t b = __ballot(v);
    if (v) {
        Q.q[(Q.head + Q.pn + Q.n + lanes_below(b)) & (JQ_CAP - 1)] = make_uint2(__float_as_uint(d), ti | (si << 16));
    }
    Q.n += (uint32_t)__popcll(b);
}

// An entry's target, proposed id and first hash slot.
__device__ inline void jq_entry(const NndDev &a, const JoinQ &Q, const uint32_t *col, uint32_t lane, uint32_t &t,
                                uint32_t &id, uint32_t &h, uint32_t &dbits) {
    const uint2 e = Q.q[(Q.head + lane) & (JQ_CAP - 1)];
    t = col[e.y & 0xFFFFu];
    id = col[e.y >> 16];
    h = fmix32(id) & (a.HP - 1);
    dbits = e.x;
}

// Start resolving the first min(n, 64) waiting entries.
__device__ inline void jq_issue(const NndDev &a, JoinQ &Q, const uint32_t *col) {
    wave_sync();
    const uint32_t lane = (uint32_t)lane_id();
    const uint32_t m = min(Q.n, (uint32_t)WAVE);
    if (lane < m) {
        uint32_t t, id, h, dbits;
        jq_entry(a, Q, col, lane, t, id, h, dbits);
        Q.g = *reinterpret_cast<const uint4 *>(a.phash + (size_t)t * a.HP + (h & ~3u));
    }
    Q.pn = m;
    Q.n -= m;
}

// Probe and propose the block in flight; returns this lane's count of
// proposals that reached a candidate slot.  Same probe order as `pooled`,
// whose first group load is the one `issue` made.
__device__ inline uint32_t jq_finish(const NndDev &a, JoinQ &Q, const uint32_t *col) {
    wave_sync();
    const uint32_t lane = (uint32_t)lane_id();
    uint32_t hitc = 0;
    if (lane < Q.pn) {
        uint32_t t, id, h, dbits;
        jq_entry(a, Q, col, lane, t, id, h, dbits);
        const uint4 g = Q.g;
        const uint32_t hit = (g.x == id ? 1u : 0u) | (g.y == id ? 2u : 0u) | (g.z == id ? 4u : 0u) |
                             (g.w == id ? 8u : 0u);
        const uint32_t emp = (g.x == ID_NONE ? 1u : 0u) | (g.y == ID_NONE ? 2u : 0u) |
                             (g.z == ID_NONE ? 4u : 0u) | (g.w == ID_NONE ? 8u : 0u);
        const uint32_t stop = (hit | emp) & (0xFu << (h & 3u));
        const bool in_pool = stop ? (hit & (stop & (0u - stop))) != 0
                                  : pooled(a, t, id, ((h & ~3u) + 4) & (a.HP - 1));
        if (!in_pool) {
            hitc = 1;
            atomicMin(reinterpret_cast<unsigned long long *>(a.cand + (size_t)t * a.C + cand_slot(t, id, a.C - 1)),
                      (unsigned long long)make_key(__uint_as_float(dbits), id));
        }
    }
    Q.head = (Q.head + Q.pn) & (JQ_CAP - 1);
    Q.pn = 0;
    wave_sync();
    return hitc;
}

// Keep at most 63 entries waiting: finish the block in flight, start the next.
__device__ inline uint32_t jq_drain(const NndDev &a, JoinQ &Q, const uint32_t *col) {
    uint32_t hits = 0;
    while (Q.n >= (uint32_t)WAVE) {
        if (Q.pn) hits += jq_finish(a, Q, col);
        jq_issue(a, Q, col);
    }
    return hits;
}

// At a node's end (fewer than 64 waiting): the waiting entries' first hash
// groups load before the block in flight is finished, so both blocks
// resolve after one wait instead of two in a row.
__device__ inline uint32_t jq_final(const NndDev &a, JoinQ &Q, const uint32_t *col) {
    wave_sync();
    const uint32_t lane = (uint32_t)lane_id();
    const uint32_t m = Q.n;
    uint4 g2 = make_uint4(0, 0, 0, 0);
    if (lane < m) {
        const uint2 e = Q.q[(Q.head + Q.pn + lane) & (JQ_CAP - 1)];
        const uint32_t t = col[e.y & 0xFFFFu], id = col[e.y >> 16];
        const uint32_t h = fmix32(id) & (a.HP - 1);
        g2 = *reinterpret_cast<const uint4 *>(a.phash + (size_t)t * a.HP + (h & ~3u));
    }
    uint32_t hits = 0;
    if (Q.pn) hits += jq_finish(a, Q, col);
    Q.g = g2;
    Q.pn = m;
    Q.n = 0;
    if (m) hits += jq_finish(a, Q, col);
    return hits;
}

// Proposals of tile pair (rt, ct) from its dot products: lane holds rows
// 16rt+4g+i, column 16ct+cl.  Each valid pair proposes both directions;
// a proposal at or below the target's worst (LDS) enters the ring.  The
// ring is drained here only when it could not hold the next row's pushes.
template <int DT>
__device__ inline void tile_proposals(const NndDev &a, JoinQ &Q, const uint32_t *col, const uint64_t *cthr,
                                      const float *cnorm, const float *cscale, uint32_t nn, uint32_t nc, int rt,
                                      int ct, floatx4 acc, uint32_t &pairs, uint32_t &props) {
    const int lane = lane_id(), g = lane >> 4, cl = lane & 15;
    const uint32_t b = (uint32_t)(ct * 16 + cl);
    const bool bok = b < nc;
    const uint32_t ib = bok ? col[b] : 0u;
    const uint64_t tb = bok ? cthr[b] : 0u;
    const float nb = bok ? cnorm[b] : 0.f;
    const float sb = (DT == DT_FP8 && bok) ? cscale[b] : 1.f;
    // one copy of the row body (and of its rare drain), not four
#pragma unroll 1
    for (int i = 0; i < 4; ++i) {
        const uint32_t ar = (uint32_t)(rt * 16 + 4 * g + i);
        // new x new: i < j once
        bool ok = bok && ar < nn && !(b < nn && !(ar < b));
        const uint32_t ia = ok ? col[ar] : 0u;
        ok = ok && ia != ib;
        bool v0 = false, v1 = false;
        float d = 0.f;
        if (ok) {
            float dot = i == 0 ? acc[0] : i == 1 ? acc[1] : i == 2 ? acc[2] : acc[3];
            if (DT == DT_FP8) dot = dot * (cscale[ar] * sb);
            d = (cnorm[ar] + nb) - 2.0f * dot;
            d = fmaxf(d, 0.0f);
            ++pairs;
            v0 = make_key(d, ib) <= cthr[ar];  // ib into ia's pool
            v1 = make_key(d, ia) <= tb;        // ia into ib's pool
        }
        jq_push(Q, v0, d, ar, b);
        jq_push(Q, v1, d, b, ar);
        // room for the next 128 pushes (the caller drains at its own point)
        if (Q.pn + Q.n > JQ_CAP - 2 * WAVE) props += jq_drain(a, Q, col);
    }
}

// Wavefronts per SIMD the join is compiled for: 4 (at most 128 VGPRs),
// except FP8 rows of 9-16 MFMA steps, whose tiles spill under that cap.
template <int DT, int TS> struct JoinWaves { static constexpr int value = (DT == DT_FP8 && TS == 16) ? 3 : 4; };

// join(), src/component_refine.cpp:127-143 / nhood::join,
// include/weavess/index.h:137-149, one wavefront per node.
// TS > 0: rows fit one K chunk of at most TS MFMA steps; operands are
// loaded to registers, RT row tiles share each column sweep, and the next
// column tile loads while the current one's proposals are made.
// TS == 0: longer rows, streamed through registers a piece at a time.
template <int DT, int TS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(JoinWaves<DT, TS>::value)))
nnd_join_kernel(NndDev a) {
    extern __shared__ __align__(16) uint8_t smem[];
    constexpr int KSTEP = JoinT<DT>::KSTEP;
    const int wid = threadIdx.x >> 6, lane = lane_id();
    const uint32_t g = blockIdx.x * (blockDim.x >> 6) + wid;
    if (g >= a.N) return;
    const uint32_t n = node_at(a, g);
    const uint32_t ncap = a.ncap;  // |new| <= S+R (2S in EFANNA's first round), |old| <= 2R
    uint8_t *my = smem + (size_t)wid * (a.lds_join_list + JQ_BYTES);
    uint32_t *nw = reinterpret_cast<uint32_t *>(my);              // nw_pow2
    uint32_t *col = nw + a.nw_pow2;                               // ncap
    uint64_t *cthr = reinterpret_cast<uint64_t *>(col + ((ncap + 1) & ~1u));
    float *cnorm = reinterpret_cast<float *>(cthr + ncap);
    float *cscale = cnorm + ncap;
    JoinQ Q{reinterpret_cast<uint2 *>(my + a.lds_join_list), 0u, 0u, 0u, make_uint4(0, 0, 0, 0)};

    const uint32_t fnc = a.fnew_cnt[n], rnc = a.rsel_new_cnt[n];
    const uint32_t foc = a.fold_cnt[n], roc = a.rsel_old_cnt[n];
    // the first 64 entries of the four lists load with the counts, not after
    // them (rows hold fst, R, L and R entries; entries past a count are
    // loaded but never used)
    const size_t bf = (size_t)n * a.fst, br = (size_t)n * a.R, bo = (size_t)n * a.L;
    const uint32_t fn0 = (uint32_t)lane < a.fst ? key_id(a.fnew[bf + lane]) : ID_NONE;
    const uint32_t rn0 = (uint32_t)lane < a.R ? a.rsel_new[br + lane] : ID_NONE;
    const uint32_t fo0 = (uint32_t)lane < a.L ? key_id(a.fold[bo + lane]) : ID_NONE;
    const uint32_t ro0 = (uint32_t)lane < a.R ? a.rsel_old[br + lane] : ID_NONE;
    const uint32_t nn = fnc + rnc;
    const uint32_t no = min(foc + roc, 2 * a.R);
    if (nn == 0) return;
    // new entry i: fnew's i-th, then rsel_new's (i - fnc)-th; sorted
    // ascending (position order == id order) in registers when they fit
    // one wavefront, else in LDS over the width in use
    if (nn <= (uint32_t)WAVE) {
        const uint32_t i = (uint32_t)lane;
        const uint32_t rs = (uint32_t)__shfl((int)rn0, (int)((i - fnc) & (WAVE - 1)));
        const uint32_t v = i < fnc ? fn0 : i < nn ? rs : ID_NONE;
        int w = 2;
        while ((uint32_t)w < nn) w <<= 1;
        const uint32_t sv = (uint32_t)wave_sort64_n((uint64_t)v, w);  // lanes from w up hold ID_NONE
        if (i < nn) col[i] = sv;
    } else {
        int pw = 2 * WAVE;  // <= nw_pow2
        while ((uint32_t)pw < nn) pw <<= 1;
        for (uint32_t i = lane; i < (uint32_t)pw; i += WAVE)
            nw[i] = i < fnc ? key_id(a.fnew[bf + i]) : i < nn ? a.rsel_new[br + (i - fnc)] : ID_NONE;
        wave_sync();
        lds_sort32(nw, pw);
        for (uint32_t i = lane; i < nn; i += WAVE) col[i] = nw[i];
    }
    // old entry j (list position nn + j): fold's j-th, then rsel_old's (j - foc)-th
    {
        const uint32_t j = (uint32_t)lane;
        const uint32_t rs = (uint32_t)__shfl((int)ro0, (int)((j - foc) & (WAVE - 1)));
        if (j < no) col[nn + j] = j < foc ? fo0 : rs;
        for (uint32_t jj = j + WAVE; jj < no; jj += WAVE)
            col[nn + jj] = jj < foc ? key_id(a.fold[bo + jj]) : a.rsel_old[br + (jj - foc)];
    }
    wave_sync();
    const uint32_t nc = nn + no;
    // each entry's threshold, norm and scale in one 16-byte record; the first
    // 128 records stay in flight while the first row tiles are requested
    // (unconditional, node 0's record for lanes past the list: a load under
    // a branch is waited for inside it)
    const uint4 rec0 = *reinterpret_cast<const uint4 *>(a.rec + ((uint32_t)lane < nc ? col[lane] : 0u));
    const uint4 rec1 =
        *reinterpret_cast<const uint4 *>(a.rec + ((uint32_t)lane + WAVE < nc ? col[lane + WAVE] : 0u));
    auto put_records = [&]() {
        auto put = [&](uint32_t i, uint4 e) {
            cthr[i] = (uint64_t)e.x | ((uint64_t)e.y << 32);
            cnorm[i] = __uint_as_float(e.z);
            if (DT == DT_FP8) cscale[i] = __uint_as_float(e.w);  // only the FP8 list has this array
        };
        // rec1 is selected outside the branch, so its load is not moved into it
        const uint4 e1 = (uint32_t)lane + WAVE < nc ? rec1 : make_uint4(0, 0, 0, 0);
        if ((uint32_t)lane < nc) put(lane, rec0);
        if ((uint32_t)lane + WAVE < nc) put(lane + WAVE, e1);
        for (uint32_t i = lane + 2 * WAVE; i < nc; i += WAVE) put(i, *reinterpret_cast<const uint4 *>(a.rec + col[i]));
        wave_sync();
    };

    const uint32_t kdim = a.kdim;
    const int nrt = (int)((nn + 15) / 16), nct = (int)((nc + 15) / 16);
    uint32_t pairs = 0, props = 0;
    if constexpr (TS > 0) {
        using V = typename JoinT<DT>::V;
        constexpr int RT = TS * (int)sizeof(V) <= 64 ? 2 : 1;
        const int T = (int)((kdim + KSTEP - 1) / KSTEP);
        // A sweep takes row tiles rt0 .. rt0+RT-1.  They are also column
        // tiles rt0 .. rt0+RT-1, so A[q] is gathered once with every listed
        // row (rows at or past nn give dot products no proposal reads), and
        // the column tiles after them stream through B, the first one
        // requested with A.  Tiles past the list load as zeros, without
        // memory accesses.
        V A[RT][TS], B[TS];
        auto load_sweep = [&](int rt0) {
#pragma unroll
            for (int q = 0; q < RT; ++q) load_tile<DT, TS>(a, col, (rt0 + q) * 16, (int)nc, T, A[q]);
            load_tile<DT, TS>(a, col, (rt0 + RT) * 16, (int)nc, T, B);
        };
        // Every load here is issued unconditionally (row 0 past the list), so
        // a wait on one tile counts the loads behind it and does not wait for
        // them.  The ring is drained right after a streamed tile's MFMAs,
        // before the next tile is requested: a drain waits for its hash
        // group loads, which would otherwise also wait for that tile.
        const floatx4 z = {0.f, 0.f, 0.f, 0.f};
        load_sweep(0);
        put_records();
        for (int rt0 = 0; rt0 < nrt; rt0 += RT) {
            if (rt0 > 0) load_sweep(rt0);
            for (int ct = rt0; ct < nct; ++ct) {
                const int j = ct - rt0;  // column tile ct is A[j] while j < RT, else B
                floatx4 acc[RT];
#pragma unroll
                for (int q = 0; q < RT; ++q) {
                    if (rt0 + q < nrt && j >= q) {
                        if (j >= RT) acc[q] = ops_mfma_k<DT, TS>(A[q], B, T, kdim, z);
                        else if (RT > 1 && j == 1) acc[q] = ops_mfma_k<DT, TS>(A[q], A[RT - 1], T, kdim, z);
                        else acc[q] = ops_mfma_k<DT, TS>(A[q], A[0], T, kdim, z);
                    }
                }
                if (j >= RT) {
                    props += jq_drain(a, Q, col);
                    // the next column tile loads while these proposals are made
                    load_tile<DT, TS>(a, col, (ct + 1) * 16, (int)nc, T, B);
                }
#pragma unroll
                for (int q = 0; q < RT; ++q)
                    if (rt0 + q < nrt && j >= q)
                        tile_proposals<DT>(a, Q, col, cthr, cnorm, cscale, nn, nc, rt0 + q, ct, acc[q], pairs,
                                           props);
            }
        }
    } else {
        put_records();
        // Longer rows, k outermost over blocks of four tiles: slots 0-1 hold
        // row tiles rt0 and rt0 + 1, slots 2-3 column tiles ct0 and ct0 + 1
        // (from rt0 + 2 on, two at a time).  Every k piece of the four tiles
        // is loaded once and meets each pair of them the join needs, so a
        // listed row is gathered once per block instead of once per tile
        // pair (for lists of up to four tiles, once).  A pair's steps are
        // accumulated in step order, as tile by tile.  Tiles past the list
        // read row 0 (loads stay unconditional); their products are never
        // used.
        using V = typename JoinT<DT>::V;
        constexpr int PS = JoinT<DT>::PIECE;
        const int T = (int)((kdim + KSTEP - 1) / KSTEP);
        const floatx4 z = {0.f, 0.f, 0.f, 0.f};
        for (int rt0 = 0; rt0 < nrt; rt0 += 2) {
            const bool r1 = rt0 + 1 < nrt;   // slot 1 is a row tile
            const bool c1 = rt0 + 1 < nct;   // slot 1 is a column tile
            for (int ct0 = rt0 + 2; ct0 == rt0 + 2 || ct0 < nct; ct0 += 2) {
                const bool first = ct0 == rt0 + 2;  // the block that also pairs slots 0-1
                const bool c2 = ct0 < nct, c3 = ct0 + 1 < nct;
                // 0: (0,0)  1: (0,1)  2: (1,1)  3: (0,2)  4: (1,2)  5: (0,3)  6: (1,3)
                floatx4 acc[7];
#pragma unroll
                for (int i = 0; i < 7; ++i) acc[i] = z;
                for (int t0 = 0; t0 < T; t0 += PS) {
                    V S0[PS], S1[PS], S2[PS], S3[PS];
                    load_piece<DT, PS>(a, col, rt0 * 16, (int)nc, T, t0, S0);
                    load_piece<DT, PS>(a, col, (rt0 + 1) * 16, (int)nc, T, t0, S1);
                    load_piece<DT, PS>(a, col, ct0 * 16, (int)nc, T, t0, S2);
                    load_piece<DT, PS>(a, col, (ct0 + 1) * 16, (int)nc, T, t0, S3);
                    const int ns = min(PS, T - t0);  // steps of this piece
                    // FP32 only: this lane's last step runs past kdim
                    bool zt = false;
                    if (DT == DT_F32 && t0 + ns == T && (kdim & (KSTEP - 1)))
                        zt = (uint32_t)(KSTEP * (T - 1) + (KSTEP / 4) * (lane >> 4)) >= kdim;
                    if (first) {
                        acc[0] = ops_mfma<DT, PS>(S0, S0, ns, acc[0], zt);
                        if (c1) acc[1] = ops_mfma<DT, PS>(S0, S1, ns, acc[1], zt);
                        if (r1) acc[2] = ops_mfma<DT, PS>(S1, S1, ns, acc[2], zt);
                    }
                    if (c2) {
                        acc[3] = ops_mfma<DT, PS>(S0, S2, ns, acc[3], zt);
                        if (r1) acc[4] = ops_mfma<DT, PS>(S1, S2, ns, acc[4], zt);
                    }
                    if (c3) {
                        acc[5] = ops_mfma<DT, PS>(S0, S3, ns, acc[5], zt);
                        if (r1) acc[6] = ops_mfma<DT, PS>(S1, S3, ns, acc[6], zt);
                    }
                }
                auto prop = [&](int rt, int ct, const floatx4 &x) {
                    tile_proposals<DT>(a, Q, col, cthr, cnorm, cscale, nn, nc, rt, ct, x, pairs, props);
                    props += jq_drain(a, Q, col);
                };
                if (first) {
                    prop(rt0, rt0, acc[0]);
                    if (c1) prop(rt0, rt0 + 1, acc[1]);
                    if (r1) prop(rt0 + 1, rt0 + 1, acc[2]);
                }
                if (c2) {
                    prop(rt0, ct0, acc[3]);
                    if (r1) prop(rt0 + 1, ct0, acc[4]);
                }
                if (c3) {
                    prop(rt0, ct0 + 1, acc[5]);
                    if (r1) prop(rt0 + 1, ct0 + 1, acc[6]);
                }
            }
        }
    }
    props += jq_drain(a, Q, col);
    props += jq_final(a, Q, col);
    pairs = wave_sum_u32(pairs);
    props = wave_sum_u32(props);
    if (lane == 0) {
        unsigned long long *c = ctr_slot(a.ctr);
        atomicAdd(&c[0], (unsigned long long)pairs);
        atomicAdd(&c[1], (unsigned long long)props);
        atomicAdd(&c[2], (unsigned long long)nc);
    }
}

// ------------------------------------------------------------------ order
// Processing order: nodes grouped by the nearest of M pivot rows (M a
// multiple of 16, pivot j = node floor(j N / M)), by squared distance over
// the first kd <= 128 FP32 elements of the base rows.  Only the kernels'
// speed depends on it (see node_at).
constexpr uint32_t ORDER_MIN_N = 32768, ORDER_MAX_PIVOTS = 256;

__device__ inline uint32_t pivot_node(uint32_t j, uint32_t N, uint32_t M) {
    return (uint32_t)((uint64_t)j * N / M);
}

__global__ void __launch_bounds__(256) nnd_pivot_norm_kernel(const float *base, uint32_t dp, uint32_t kd, uint32_t N,
                                                             uint32_t M, float *pnorm) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    const float *row = base + (size_t)pivot_node(j, N, M) * dp;
    float s = 0.f;
    for (uint32_t e = 0; e < kd; ++e) s += row[e] * row[e];
    pnorm[j] = s;
}

// One wavefront per 16 nodes: |p|^2 - 2 x.p for 16 pivots at a time on
// MFMA (16x16x4 f32, operands as in the join's FP32 tiles); each lane keeps
// the smallest of its column across the pivot tiles, and one reduction over
// the 16 lanes of a row group gives each node its nearest (ties: the lower
// pivot).  Loads are issued without branches: steps past T repeat step T-1
// (not used) and elements at or past kd read the row's first vector and are
// zeroed at the MFMA.
__device__ inline void pivot_tile_ops(const float *base, uint32_t dp, uint32_t kd, uint32_t row_node, int T,
                                      uint4 ops[8]) {
    const int g = lane_id() >> 4;
    const float *row = base + (size_t)row_node * dp;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint32_t e0 = (uint32_t)(16 * min(j, T - 1) + 4 * g);
        ops[j] = *reinterpret_cast<const uint4 *>(row + (e0 < kd ? e0 : 0u));
    }
}

__global__ void __launch_bounds__(256) nnd_assign_kernel(const float *base, uint32_t dp, uint32_t kd, uint32_t N,
                                                         uint32_t M, const float *pnorm, uint32_t *label) {
    const int lane = lane_id(), r = lane & 15, g = lane >> 4;
    const uint32_t n0 = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 16;
    if (n0 >= N) return;
    const int T = (int)((kd + 15) / 16);  // steps of 16 elements, 4 MFMAs each
    const bool ztail = (uint32_t)(16 * (T - 1) + 4 * g) >= kd;
    uint4 A[8], B[8];
    pivot_tile_ops(base, dp, kd, min(n0 + (uint32_t)r, N - 1), T, A);  // rows past N: labels not written
    float bv[4];
    uint32_t bi[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        bv[i] = INFINITY;
        bi[i] = 0;
    }
    for (uint32_t t = 0; t < M / 16; ++t) {
        const uint32_t col = t * 16 + (uint32_t)r;  // output: lane holds column r of rows 4g + i
        pivot_tile_ops(base, dp, kd, pivot_node(col, N, M), T, B);
        const floatx4 acc = ops_mfma<DT_F32, 8>(A, B, T, floatx4{0.f, 0.f, 0.f, 0.f}, ztail);
        const float pn = pnorm[col];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float v = pn - 2.0f * acc[i];
            if (v < bv[i]) {  // columns ascend, so an equal value keeps the lower pivot
                bv[i] = v;
                bi[i] = col;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {  // within the 16 lanes of row group g
            const float ov = __shfl_xor(bv[i], o);
            const uint32_t ok = (uint32_t)__shfl_xor((int)bi[i], o);
            if (ov < bv[i] || (ov == bv[i] && ok < bi[i])) {
                bv[i] = ov;
                bi[i] = ok;
            }
        }
    }
    if (r == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t n = n0 + 4 * (uint32_t)g + (uint32_t)i;
            if (n < N) label[n] = bi[i];
        }
    }
}

// Counting sort of the nodes by label: per-label counts, their exclusive
// scan, then each workgroup ranks its nodes in LDS and claims one range per
// label.  The order within a label is whatever the atomics give.
__global__ void __launch_bounds__(256) nnd_label_hist_kernel(const uint32_t *label, uint32_t N, uint32_t M,
                                                             uint32_t *counts) {
    __shared__ uint32_t h[ORDER_MAX_PIVOTS];
    for (uint32_t i = threadIdx.x; i < M; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x)
        atomicAdd(&h[label[n]], 1u);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < M; i += blockDim.x)
        if (h[i]) atomicAdd(&counts[i], h[i]);
}

__global__ void nnd_label_scan_kernel(const uint32_t *counts, uint32_t M, uint32_t *offs) {
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (uint32_t b = 0; b < M; ++b) {
            offs[b] = acc;
            acc += counts[b];
        }
    }
}

constexpr uint32_t ORDER_CHUNK = 4096;  // nodes a workgroup of the scatter ranks in LDS

__global__ void __launch_bounds__(256) nnd_label_scatter_kernel(const uint32_t *label, uint32_t N, uint32_t M,
                                                                uint32_t *offs, uint32_t *order) {
    __shared__ uint32_t h[ORDER_MAX_PIVOTS], start[ORDER_MAX_PIVOTS];
    const uint32_t c0 = blockIdx.x * ORDER_CHUNK, c1 = min(c0 + ORDER_CHUNK, N);
    for (uint32_t i = threadIdx.x; i < M; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (uint32_t n = c0 + threadIdx.x; n < c1; n += blockDim.x) atomicAdd(&h[label[n]], 1u);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < M; i += blockDim.x) {
        start[i] = h[i] ? atomicAdd(&offs[i], h[i]) : 0u;
        h[i] = 0;
    }
    __syncthreads();
    for (uint32_t n = c0 + threadIdx.x; n < c1; n += blockDim.x) {
        const uint32_t b_ = label[n];
        order[start[b_] + atomicAdd(&h[b_], 1u)] = n;
    }
}

// ------------------------------------------------------------------ pool hash
// Rebuild phash[n] from the pool (load factor <= 1/2, so probes are short
// and every lookup terminates at an empty slot).
__global__ void __launch_bounds__(256) nnd_hash_kernel(NndDev a) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int wid = threadIdx.x >> 6, lane = lane_id();
    const uint32_t n = blockIdx.x * (blockDim.x >> 6) + wid;
    if (n >= a.N) return;
    uint32_t *hk = reinterpret_cast<uint32_t *>(smem) + (size_t)wid * a.HP;
    const uint32_t mask = a.HP - 1;
    for (uint32_t i = lane; i < a.HP; i += WAVE) hk[i] = ID_NONE;
    wave_sync();
    const uint32_t cnt = a.pcnt[n];
    for (uint32_t i = lane; i < cnt; i += WAVE) {
        const uint32_t id = key_id(a.pool[(size_t)n * a.L + i]);
        uint32_t h = fmix32(id) & mask;
        while (atomicCAS(&hk[h], ID_NONE, id) != ID_NONE) h = (h + 1) & mask;
    }
    wave_sync();
    uint32_t *dst = a.phash + (size_t)n * a.HP;
    for (uint32_t i = lane; i < a.HP; i += WAVE) dst[i] = hk[i];
}

// ------------------------------------------------------------------ merge
__host__ __device__ inline size_t merge_wave_lds(uint32_t up2, uint32_t hp2, uint32_t L) {
    return ((size_t)up2 * 8 + (size_t)L * 8 + (size_t)hp2 * 5 + L + 16 + 15) & ~(size_t)15;
}
// Fold each node's slot winners into its pool (index.h:120-135 applied to
// the whole batch at once; see header comment), then compute the next M and,
// when `rehash`, rebuild the node's pool-id hash for the next join.  All of
// the node's global loads (C slots, L pool keys and flags; C, L <= 1024, so
// at most 16 per lane each) are issued before any of them is used.
// MR: key registers per lane, >= C / 64 and >= L / 64 (at most 16).
template <int MR>
__global__ void __launch_bounds__(256) nnd_merge_kernel(NndDev a, uint32_t up2, uint32_t hp2, int rehash) {
    constexpr int MERGE_REGS = MR;
    extern __shared__ __align__(16) uint8_t smem[];
    const int wid = threadIdx.x >> 6, lane = lane_id();
    const uint32_t n = blockIdx.x * (blockDim.x >> 6) + wid;
    if (n >= a.N) return;
    uint8_t *my = smem + (size_t)wid * merge_wave_lds(up2, hp2, a.L);
    uint64_t *u = reinterpret_cast<uint64_t *>(my);   // pool keys, then the winners
    uint64_t *w = u + up2;                             // the merged, truncated pool
    uint32_t *hk = reinterpret_cast<uint32_t *>(w + a.L);
    uint8_t *hf = reinterpret_cast<uint8_t *>(hk + hp2);
    uint8_t *nf = hf + hp2;
    const uint32_t hmask = hp2 - 1;
    const uint32_t cnt = a.pcnt[n];
    const uint64_t *pr = a.pool + (size_t)n * a.L;
    const uint8_t *pf = a.pflag + (size_t)n * a.L;
    uint64_t *cs = a.cand + (size_t)n * a.C;
    const uint32_t cper = a.C / WAVE;
    uint64_t ck[MERGE_REGS], pk[MERGE_REGS];
    uint8_t pfl[MERGE_REGS];
#pragma unroll
    for (int r = 0; r < MERGE_REGS; ++r) {
        const uint32_t i = lane + r * WAVE;
        ck[r] = (uint32_t)r < cper ? cs[i] : KEY_EMPTY;
        // loaded with the count, not after it (a row holds L slots)
        const uint64_t k = i < a.L ? pr[i] : KEY_EMPTY;
        const uint8_t f = i < a.L ? pf[i] : (uint8_t)0;
        pk[r] = i < cnt ? k : KEY_EMPTY;
        pfl[r] = i < cnt ? f : (uint8_t)0;
    }
    // Slot winners are never ids already in the pool: the join probed each
    // proposal against its target's pool-id hash before the slot update.
    // Pools of at most 64 with at most 64 winners (the register path below)
    // therefore need no hash of the old pool; the general path builds one
    // for the flags it carries over.
    uint32_t nwin = 0;
#pragma unroll
    for (int r = 0; r < MERGE_REGS; ++r)
        if ((uint32_t)r < cper) nwin += (uint32_t)__popcll(__ballot(ck[r] != KEY_EMPTY));
    const bool in_regs = a.L <= (uint32_t)WAVE && nwin <= (uint32_t)WAVE;
    if (!in_regs) {
        for (uint32_t i = lane; i < hp2; i += WAVE) hk[i] = ID_NONE;
        wave_sync();
#pragma unroll
        for (int r = 0; r < MERGE_REGS; ++r) {
            const uint32_t i = lane + r * WAVE;
            if (i >= cnt) continue;
            u[i] = pk[r];
            uint32_t h = fmix32(key_id(pk[r])) & hmask;
            while (atomicCAS(&hk[h], ID_NONE, key_id(pk[r])) != ID_NONE) h = (h + 1) & hmask;
            hf[h] = pfl[r];
        }
        wave_sync();
    }
    // reset every slot for the next join with whole-line stores (resetting
    // only the filled ones leaves scattered partial-line writes)
#pragma unroll
    for (int r = 0; r < MERGE_REGS; ++r)
        if ((uint32_t)r < cper) cs[lane + r * WAVE] = KEY_EMPTY;
    // winners (not already pooled), compacted behind the pool keys by ballot
    // (slot hashing gives each id one slot, so they are distinct)
    uint32_t nw = 0;
#pragma unroll
    for (int r = 0; r < MERGE_REGS; ++r) {
        if ((uint32_t)r >= cper) continue;  // uniform; `break` would keep the loop rolled
        const uint64_t k = ck[r];
        bool take = false;
        if (k != KEY_EMPTY) {
            const uint32_t id = key_id(k);
            uint32_t h = fmix32(id) & hmask;
            take = true;
            if (!in_regs) {
                for (;;) {
                    const uint32_t e = hk[h];
                    if (e == id) { take = false; break; }
                    if (e == ID_NONE) break;
                    h = (h + 1) & hmask;
                }
            }
        }
        const uint64_t bm = __ballot(take);
        if (take) u[cnt + nw + lanes_below(bm)] = k;
        nw += (uint32_t)__popcll(bm);
    }
    if (a.L <= (uint32_t)WAVE && nw <= (uint32_t)WAVE) {
        // Pools of at most 64 with at most 64 winners, in registers: lane i
        // holds pool key i (its flag in bit 31) and, after a sort across the
        // wavefront, winner i (tagged new in bit 63, flagged).  Comparing
        // each pool key with the winner of the mirrored lane leaves the 64
        // smallest of both runs as a bitonic sequence, and six merge steps
        // sort it: the new pool is its first L, the same keys in the same
        // order as the merge by rank below, with the same flags.
        constexpr uint64_t NEWB = 1ull << 63, FLB = 1ull << 31, CM = ~(NEWB | FLB);
        wave_sync();
        int pw = 2;
        while ((uint32_t)pw < nw) pw <<= 1;
        // lanes from pw up hold KEY_EMPTY before and after the sort
        uint64_t wk = wave_sort64_n((uint32_t)lane < nw ? u[cnt + lane] : KEY_EMPTY, pw);
        if (wk != KEY_EMPTY) wk |= NEWB | FLB;
        const uint64_t pkey = (uint32_t)lane < cnt ? (pk[0] | (pfl[0] ? FLB : 0ull)) : KEY_EMPTY;
        const uint64_t rw = shfl64(wk, WAVE - 1 - lane);
        uint64_t x = (pkey & CM) < (rw & CM) ? pkey : rw;
#pragma unroll
        for (int j = WAVE / 2; j > 0; j >>= 1) {
            const uint64_t o = shfl_xor64(x, j);
            const bool below = (x & CM) < (o & CM);
            x = ((lane & j) == 0) == below ? x : o;
        }
        const uint32_t keep = min(cnt + nw, a.L);
        const bool in = (uint32_t)lane < keep;
        const uint64_t k = in ? (x & CM) : KEY_EMPTY;
        const bool f = in && (x & FLB) != 0;
        if ((uint32_t)lane < a.L) {
            a.pool[(size_t)n * a.L + lane] = k;
            a.pflag[(size_t)n * a.L + lane] = f ? 1 : 0;
        }
        const uint32_t fresh = (uint32_t)__popcll(__ballot(in && (x & NEWB) != 0));
        // compute_M on the new flags (one ballot: the pool holds <= 64)
        uint32_t m = 0;
        if (a.S > 0) {
            const uint32_t maxl = min(a.M[n] + a.S, keep);
            uint64_t b = __ballot((uint32_t)lane < maxl && f);
            if ((uint32_t)__popcll(b) >= a.S) {
                for (uint32_t i = 1; i < a.S; ++i) b &= b - 1;
                m = (uint32_t)__ffsll((unsigned long long)b);  // position + 1
            } else {
                m = maxl;
            }
        }
        const uint64_t last = keep ? readlane64(k, (int)keep - 1) : KEY_EMPTY;
        if (lane == 0) {
            a.pcnt[n] = keep;
            a.rec[n].thr = last;
            a.worst[n] = keep ? key_dist(last) : -1.0f;
            a.M[n] = m;
            atomicAdd(&ctr_slot(a.ctr)[3], (unsigned long long)fresh);
        }
        if (rehash) {
            wave_sync();
            for (uint32_t i = lane; i < hp2; i += WAVE) hk[i] = ID_NONE;
            wave_sync();
            if (in) {
                const uint32_t id = key_id(k);
                uint32_t h = fmix32(id) & hmask;
                while (atomicCAS(&hk[h], ID_NONE, id) != ID_NONE) h = (h + 1) & hmask;
            }
            wave_sync();
            uint32_t *dst = a.phash + (size_t)n * a.HP;
            for (uint32_t i = lane; i < hp2; i += WAVE) dst[i] = hk[i];
        }
        return;
    }
    // sort the winners only (the pool part is sorted), then merge the two runs
    // by rank: a key's slot is its index plus the other run's keys below it
    if (nw > 0) {
        uint32_t pw = 64;
        while (pw < nw) pw <<= 1;  // <= C, so cnt + pw <= L + C <= up2
        wave_sync();
        for (uint32_t i = nw + lane; i < pw; i += WAVE) u[cnt + i] = KEY_EMPTY;
        wave_sync();
        lds_sort64(u + cnt, (int)pw);
    }
    wave_sync();
    const uint64_t *bw = u + cnt;
    for (uint32_t i = lane; i < cnt; i += WAVE) {
        const uint32_t p = i + (uint32_t)lower_bound64(bw, (int)nw, u[i]);
        if (p < a.L) w[p] = u[i];
    }
    for (uint32_t j = lane; j < nw; j += WAVE) {
        const uint32_t p = j + (uint32_t)lower_bound64(u, (int)cnt, bw[j]);
        if (p < a.L) w[p] = bw[j];
    }
    wave_sync();
    const uint32_t keep = min(cnt + nw, a.L);
    uint32_t fresh = 0;
    for (uint32_t i = lane; i < a.L; i += WAVE) {
        const uint64_t k = i < keep ? w[i] : KEY_EMPTY;
        uint8_t f = 0;
        if (i < keep) {
            const uint32_t id = key_id(k);
            uint32_t h = fmix32(id) & hmask;
            f = 1;
            for (;;) {
                const uint32_t e = hk[h];
                if (e == id) { f = hf[h]; break; }
                if (e == ID_NONE) { ++fresh; break; }
                h = (h + 1) & hmask;
            }
        }
        a.pool[(size_t)n * a.L + i] = k;
        a.pflag[(size_t)n * a.L + i] = f;
        nf[i] = f;
    }
    wave_sync();
    const uint32_t m = compute_M(nf, a.M[n], a.S, keep);
    fresh = wave_sum_u32(fresh);
    if (lane == 0) {
        a.pcnt[n] = keep;
        a.rec[n].thr = keep ? w[keep - 1] : KEY_EMPTY;
        a.worst[n] = keep ? key_dist(w[keep - 1]) : -1.0f;
        a.M[n] = m;
        atomicAdd(&ctr_slot(a.ctr)[3], (unsigned long long)fresh);
    }
    if (rehash) {
        // the new pool's id hash (hp2 == HP: load factor <= 1/2)
        for (uint32_t i = lane; i < hp2; i += WAVE) hk[i] = ID_NONE;
        wave_sync();
        for (uint32_t i = lane; i < keep; i += WAVE) {
            const uint32_t id = key_id(w[i]);
            uint32_t h = fmix32(id) & hmask;
            while (atomicCAS(&hk[h], ID_NONE, id) != ID_NONE) h = (h + 1) & hmask;
        }
        wave_sync();
        uint32_t *dst = a.phash + (size_t)n * a.HP;
        for (uint32_t i = lane; i < hp2; i += WAVE) dst[i] = hk[i];
    }
}

// ------------------------------------------------------------------ finalize
// RefineInner :22-74 + PruneNaive: exact FP32 build-order distances for the
// first `rerank` pool entries, ascending, first K kept.
template <int CPL>
__global__ void __launch_bounds__(256) nnd_finalize_kernel(NndDev a, uint32_t rp2) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int wid = threadIdx.x >> 6, lane = lane_id();
    const uint32_t g = blockIdx.x * (blockDim.x >> 6) + wid;
    if (g >= a.N) return;
    const uint32_t n = node_at(a, g);
    uint8_t *my = smem + (size_t)wid * (rp2 * 12 + 16);
    uint64_t *keys = reinterpret_cast<uint64_t *>(my);
    uint32_t *ids = reinterpret_cast<uint32_t *>(keys + rp2);
    const uint32_t cnt = a.pcnt[n];
    const uint32_t rr = min(cnt, a.rerank);
    for (uint32_t i = lane; i < rr; i += WAVE) ids[i] = key_id(a.pool[(size_t)n * a.L + i]);
    wave_sync();
    float4 qv[CPL];
    {
        const float4 *qp = reinterpret_cast<const float4 *>(a.base + (size_t)n * a.dp);
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            const int ch = (lane & 15) + 16 * c;
            qv[c] = ch < (int)(a.dp / 4) ? qp[ch] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
    // long rows one row group at a time, as in nnd_init_kernel
    rows_dist<CPL, (CPL >= 8 ? 1 : 4)>(a.base, a.dp, a.dp / 4, qv, ids, (int)rr, keys);
    for (uint32_t i = rr + lane; i < rp2; i += WAVE) keys[i] = KEY_EMPTY;
    wave_sync();
    lds_sort64(keys, (int)rp2);
    const uint32_t deg = min(a.K, rr);
    for (uint32_t i = lane; i < a.K; i += WAVE) {
        const bool ok = i < deg;
        a.out_ids[(size_t)n * a.K + i] = ok ? key_id(keys[i]) : ID_NONE;
        a.out_dist[(size_t)n * a.K + i] = ok ? key_dist(keys[i]) : -1.0f;
    }
    if (lane == 0) a.out_deg[n] = deg;
}

// ------------------------------------------------------------------ host
uint32_t pow2_at_least(uint32_t v, uint32_t lo) {
    uint32_t p = lo;
    while (p < v) p <<= 1;
    return p;
}

struct Bufs {
    std::vector<void *> ptrs;
    ~Bufs() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <class T> hipError_t get(T **p, size_t n) {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(n * sizeof(T), 16));
        if (e == hipSuccess) {
            ptrs.push_back(q);
            *p = static_cast<T *>(q);
        }
        return e;
    }
};

constexpr int WPB = 1;       // wavefronts (nodes) per workgroup: one, as in the join (four were 1-8 % slower in every stage)
constexpr int JOIN_WPB = 1;  // one node a workgroup: a finished wavefront frees its slot at once

// The join kernel for this element type: register operands of at most ts
// MFMA steps (rows of at most 512 bytes), or rows streamed through
// registers a piece at a time (ts == 0).
using JoinFn = void (*)(NndDev);
template <int DT>
JoinFn join_kernel(int ts) {
    if (ts == 4) return nnd_join_kernel<DT, 4>;
    if (ts == 8) return nnd_join_kernel<DT, 8>;
    if constexpr (DT == DT_FP8) {
        if (ts == 16) return nnd_join_kernel<DT, 16>;
    }
    return nnd_join_kernel<DT, 0>;
}

hipError_t scan_u32(const uint32_t *in, uint32_t n, uint32_t *out, uint32_t *bsum, hipStream_t s) {
    const uint32_t nb = (n + SCAN_B - 1) / SCAN_B;
    hipLaunchKernelGGL(scan_sums_kernel, dim3(nb), dim3(SCAN_T), 0, s, in, n, bsum);
    hipLaunchKernelGGL(scan_bsum_kernel, dim3(1), dim3(SCAN_T), 0, s, bsum, nb);
    hipLaunchKernelGGL(scan_final_kernel, dim3(nb), dim3(SCAN_T), 0, s, in, n, bsum, out);
    return hipGetLastError();
}

template <class K>
hipError_t set_lds(K kernel, size_t bytes) {
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

#define NND_HIP(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) {                                                              \
            char m_[256];                                                                    \
            snprintf(m_, sizeof m_, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                     __LINE__);                                                              \
            return index_set_error(ix, e_ == hipErrorOutOfMemory ? WV_ENOMEM : WV_EHIP, m_); \
        }                                                                                    \
    } while (0)

template <int CPL>
hipError_t launch_init(const NndDev &a, const uint32_t *ii, const uint32_t *ic, uint32_t blocks, hipStream_t s) {
    hipLaunchKernelGGL((nnd_init_kernel<CPL, DT_F32>), dim3(blocks), dim3(64 * WPB), WPB * (256 + 512), s, a, ii,
                       ic);
    return hipGetLastError();
}
template <int DT>
hipError_t launch_init_lp(const NndDev &a, const uint32_t *ii, const uint32_t *ic, uint32_t blocks, hipStream_t s) {
    hipLaunchKernelGGL((nnd_init_kernel<1, DT>), dim3(blocks), dim3(64 * WPB), WPB * (256 + 512), s, a, ii, ic);
    return hipGetLastError();
}
template <int CPL>
hipError_t launch_init_rows(const NndDev &a, const uint32_t *ii, const uint32_t *ic, uint32_t stride, uint32_t blocks,
                            hipStream_t s) {
    const uint32_t P = pow2_at_least(stride, 64);
    const size_t lds = (size_t)WPB * (256 + 512 + (size_t)P * 8);
    hipError_t e = set_lds(nnd_init_rows_kernel<CPL>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(nnd_init_rows_kernel<CPL>, dim3(blocks), dim3(64 * WPB), lds, s, a, ii, ic, stride, P);
    return hipGetLastError();
}
template <int CPL>
hipError_t launch_final(const NndDev &a, uint32_t rp2, uint32_t blocks, hipStream_t s) {
    const size_t lds = (size_t)WPB * (rp2 * 12 + 16);
    hipError_t e = set_lds(nnd_finalize_kernel<CPL>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(nnd_finalize_kernel<CPL>, dim3(blocks), dim3(64 * WPB), lds, s, a, rp2);
    return hipGetLastError();
}

}  // namespace

int nndescent_build(wv_index *ix, const wv_nnd_params *p, const uint32_t *init_ids,
                    const uint32_t *init_cnt, wv_build_stats *stats, uint32_t *graph_out,
                    float *gdist_out, uint32_t *deg_out, bool efanna, uint32_t init_stride,
                    const NndMonitor *monitor) {
    const IndexView v = index_view(ix);
    const hipStream_t s = v.stream;
    const uint32_t N = v.N;
    if (p->dtype > 2) return index_set_error(ix, WV_EINVAL, "dtype must be WV_DTYPE_F32, _BF16 or _FP8");
    const int dt = (int)p->dtype;
    const uint32_t L = p->L, S = p->S, R = p->R, K = p->K;
    const uint32_t nch = v.dp / 4;
    const int cpl = nch <= 16 ? 1 : nch <= 32 ? 2 : nch <= 64 ? 4 : nch <= 128 ? 8 : nch <= 256 ? 16 : 0;
    if (!cpl) return index_set_error(ix, WV_ENOTSUP, "dimension above 1024");
    uint32_t C = p->cand_slots ? p->cand_slots : std::min<uint32_t>(1024, pow2_at_least(4 * L, 64));
    if ((C & (C - 1)) || C < 64 || C > 1024)
        return index_set_error(ix, WV_EINVAL, "cand_slots must be a power of two in [64, 1024]");
    const uint32_t rer = std::min(L, std::max(K, (p->rerank_mult ? p->rerank_mult : (dt ? 2u : 1u)) * K));

    Bufs B;
    NndDev a{};
    a.base = v.base;
    a.dp = v.dp;
    a.N = N; a.L = L; a.S = S; a.R = R; a.K = K; a.C = C; a.rerank = rer;
    a.seed = p->seed;
    uint8_t *xq = nullptr;
    NND_HIP(B.get(&a.rec, N));
    if (dt == DT_F32) {
        a.xq = reinterpret_cast<const uint8_t *>(v.base);
        a.xq_stride = v.dp * 4;
        a.kdim = v.dp;
    } else {
        a.kdim = (v.dp + 31) & ~31u;
        a.xq_stride = a.kdim * (dt == DT_BF16 ? 2 : 1);
        NND_HIP(B.get(&xq, (size_t)N * a.xq_stride));
        NND_HIP(hipMemsetAsync(xq, 0, (size_t)N * a.xq_stride, s));
        a.xq = xq;
    }
    NND_HIP(B.get(&a.pool, (size_t)N * L));
    NND_HIP(B.get(&a.pflag, (size_t)N * L));
    NND_HIP(B.get(&a.pcnt, N));
    NND_HIP(B.get(&a.M, N));
    NND_HIP(B.get(&a.worst, N));
    a.fst = efanna ? 2 * S : S;
    NND_HIP(B.get(&a.fnew, (size_t)N * a.fst));
    NND_HIP(B.get(&a.fold, (size_t)N * L));
    NND_HIP(B.get(&a.fnew_cnt, N));
    NND_HIP(B.get(&a.fold_cnt, N));
    NND_HIP(B.get(&a.rcnt_new, N));
    NND_HIP(B.get(&a.rcnt_old, N));
    NND_HIP(B.get(&a.rpos_new, (size_t)N * a.fst));
    NND_HIP(B.get(&a.rpos_old, (size_t)N * L));
    NND_HIP(B.get(&a.roff_new, (size_t)N + 1));
    NND_HIP(B.get(&a.roff_old, (size_t)N + 1));
    NND_HIP(B.get(&a.rv_new, (size_t)N * L));
    NND_HIP(B.get(&a.rv_old, (size_t)N * L));
    NND_HIP(B.get(&a.rsel_new, (size_t)N * std::max(R, 1u)));
    NND_HIP(B.get(&a.rsel_old, (size_t)N * std::max(R, 1u)));
    NND_HIP(B.get(&a.rsel_new_cnt, N));
    NND_HIP(B.get(&a.rsel_old_cnt, N));
    NND_HIP(B.get(&a.cand, (size_t)N * C));
    a.HP = pow2_at_least(2 * L, 64);
    NND_HIP(B.get(&a.phash, (size_t)N * a.HP));
    const size_t hash_lds = (size_t)WPB * a.HP * 4;
    NND_HIP(B.get(&a.ctr, (size_t)CTR_SLOTS * 4));
    uint32_t *bsum = nullptr;
    NND_HIP(B.get(&bsum, (N + SCAN_B - 1) / SCAN_B + 1));
    a.out_ids = graph_out;
    a.out_dist = gdist_out;
    a.out_deg = deg_out;

    const uint32_t ncap = efanna ? std::max(S + 3 * R, 2 * S) : S + 3 * R;
    a.ncap = ncap;
    a.nw_pow2 = pow2_at_least(efanna ? std::max(S + R, 2 * S) : S + R, 64);
    // list area: nw ids, col ids, cthr keys, cnorm, and FP8 row scales
    a.lds_join_list = (uint32_t)(((a.nw_pow2 * 4 + ((ncap + 1) & ~1u) * 4 + ncap * (dt == DT_FP8 ? 16 : 12)) + 15) &
                                 ~15u);
    int join_ts = 0;
    {
        // MFMA steps of a row (16 FP32 or 32 BF16/FP8 elements each); rows of
        // at most one 512-byte chunk keep their operands in registers
        const uint32_t es = dt == DT_F32 ? 4 : (dt == DT_BF16 ? 2 : 1);
        const uint32_t kc = 512 / es, kstep = dt == DT_F32 ? 16 : 32;
        const uint32_t steps = (a.kdim + kstep - 1) / kstep;
        join_ts = a.kdim > kc ? 0 : steps <= 4 ? 4 : steps <= 8 ? 8 : 16;
    }
    // list area and proposal ring; the ring's 16-bit list positions hold any
    // ncap these 160 KB allow
    const size_t join_lds = (size_t)JOIN_WPB * (a.lds_join_list + JQ_BYTES);
    if (join_lds > 160 * 1024)
        return index_set_error(ix, WV_ENOTSUP, "S/R too large for the join's LDS lists");
    const uint32_t up2 = pow2_at_least(L + C, 64), hp2 = pow2_at_least(2 * L, 64);
    const size_t merge_lds = (size_t)WPB * merge_wave_lds(up2, hp2, L);
    if (merge_lds > 160 * 1024) return index_set_error(ix, WV_ENOTSUP, "L + cand_slots too large for LDS");
    const uint32_t rp2 = pow2_at_least(R, 64);
    const size_t sel_lds = (size_t)WPB * (rp2 * 4 + 256 * 4 + 16);
    const uint32_t fp2 = pow2_at_least(rer, 64);

    const uint32_t blocks = (N + WPB - 1) / WPB;
    hipEvent_t ev[8];
    for (auto &e : ev) NND_HIP(hipEventCreate(&e));
    struct EvGuard {
        hipEvent_t *e;
        ~EvGuard() {
            for (int i = 0; i < 8; ++i) (void)hipEventDestroy(e[i]);
        }
    } evg{ev};
    double t_sample = 0, t_join = 0, t_merge = 0;

    NND_HIP(hipEventRecord(ev[0], s));
    hipLaunchKernelGGL(fill_u64_kernel, dim3(2048), dim3(256), 0, s, a.cand, KEY_EMPTY, (size_t)N * C);
    hipLaunchKernelGGL(fill_u64_kernel, dim3(16), dim3(256), 0, s, reinterpret_cast<uint64_t *>(a.ctr), 0ull,
                       (size_t)CTR_SLOTS * 4);
    {
        const uint32_t pb = (N * 16 + 255) / 256;
        if (dt == DT_F32)
            hipLaunchKernelGGL(nnd_prep_kernel<DT_F32>, dim3(pb), dim3(256), 0, s, a, xq);
        else if (dt == DT_BF16)
            hipLaunchKernelGGL(nnd_prep_kernel<DT_BF16>, dim3(pb), dim3(256), 0, s, a, xq);
        else
            hipLaunchKernelGGL(nnd_prep_kernel<DT_FP8>, dim3(pb), dim3(256), 0, s, a, xq);
        NND_HIP(hipGetLastError());
    }
    // the processing order (speed only; the graph is the same without it)
    if (p->order == 2 || (p->order == 0 && N >= ORDER_MIN_N)) {
        uint32_t M = ORDER_MAX_PIVOTS;
        while (M > 16 && (size_t)M * 128 > N) M >>= 1;  // at least about 128 nodes a pivot
        const uint32_t kd = std::min(v.dp, 128u);
        uint32_t *label = nullptr, *counts = nullptr, *offs = nullptr, *order = nullptr;
        float *pnorm = nullptr;
        NND_HIP(B.get(&label, N));
        NND_HIP(B.get(&order, N));
        NND_HIP(B.get(&counts, M));
        NND_HIP(B.get(&offs, M));
        NND_HIP(B.get(&pnorm, M));
        hipLaunchKernelGGL(nnd_pivot_norm_kernel, dim3((M + 255) / 256), dim3(256), 0, s, v.base, v.dp, kd, N, M,
                           pnorm);
        const uint32_t aw = (N + 15) / 16;  // wavefronts of 16 nodes
        hipLaunchKernelGGL(nnd_assign_kernel, dim3((aw + 3) / 4), dim3(256), 0, s, v.base, v.dp, kd, N, M, pnorm,
                           label);
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(256), 0, s, counts, 0u, (size_t)M);
        hipLaunchKernelGGL(nnd_label_hist_kernel, dim3(std::min<uint32_t>(256, (N + 255) / 256)), dim3(256), 0, s,
                           label, N, M, counts);
        hipLaunchKernelGGL(nnd_label_scan_kernel, dim3(1), dim3(64), 0, s, counts, M, offs);
        hipLaunchKernelGGL(nnd_label_scatter_kernel, dim3((N + ORDER_CHUNK - 1) / ORDER_CHUNK), dim3(256), 0, s, label,
                           N, M, offs, order);
        NND_HIP(hipGetLastError());
        a.order = order;
    }
    uint32_t *d_init = nullptr, *d_icnt = nullptr;
    const uint32_t ist = efanna && init_ids ? init_stride : S;  // ids per initial row
    if (init_ids) {
        NND_HIP(B.get(&d_init, (size_t)N * ist));
        NND_HIP(hipMemcpyAsync(d_init, init_ids, (size_t)N * ist * 4, hipMemcpyHostToDevice, s));
        if (init_cnt) {
            NND_HIP(B.get(&d_icnt, N));
            NND_HIP(hipMemcpyAsync(d_icnt, init_cnt, (size_t)N * 4, hipMemcpyHostToDevice, s));
        }
    }
    if (efanna && init_ids) {
        switch (cpl) {
        case 1: NND_HIP(launch_init_rows<1>(a, d_init, d_icnt, ist, blocks, s)); break;
        case 2: NND_HIP(launch_init_rows<2>(a, d_init, d_icnt, ist, blocks, s)); break;
        case 4: NND_HIP(launch_init_rows<4>(a, d_init, d_icnt, ist, blocks, s)); break;
        case 8: NND_HIP(launch_init_rows<8>(a, d_init, d_icnt, ist, blocks, s)); break;
        default: NND_HIP(launch_init_rows<16>(a, d_init, d_icnt, ist, blocks, s)); break;
        }
    } else if (dt == DT_BF16) {
        NND_HIP(launch_init_lp<DT_BF16>(a, d_init, d_icnt, blocks, s));
    } else if (dt == DT_FP8) {
        NND_HIP(launch_init_lp<DT_FP8>(a, d_init, d_icnt, blocks, s));
    } else {
        switch (cpl) {
        case 1: NND_HIP(launch_init<1>(a, d_init, d_icnt, blocks, s)); break;
        case 2: NND_HIP(launch_init<2>(a, d_init, d_icnt, blocks, s)); break;
        case 4: NND_HIP(launch_init<4>(a, d_init, d_icnt, blocks, s)); break;
        case 8: NND_HIP(launch_init<8>(a, d_init, d_icnt, blocks, s)); break;
        default: NND_HIP(launch_init<16>(a, d_init, d_icnt, blocks, s)); break;
        }
    }
    NND_HIP(set_lds(nnd_hash_kernel, hash_lds));
    hipLaunchKernelGGL(nnd_hash_kernel, dim3(blocks), dim3(64 * WPB), hash_lds, s, a);
    NND_HIP(hipGetLastError());
    NND_HIP(hipEventRecord(ev[1], s));
    NND_HIP(hipEventSynchronize(ev[1]));
    float ms_init = 0.f;
    NND_HIP(hipEventElapsedTime(&ms_init, ev[0], ev[1]));

    const JoinFn join_fn = dt == DT_F32 ? join_kernel<DT_F32>(join_ts)
                         : dt == DT_BF16 ? join_kernel<DT_BF16>(join_ts) : join_kernel<DT_FP8>(join_ts);
    NND_HIP(set_lds(join_fn, join_lds));
    // merge: key registers per lane for max(C, L) keys
    const uint32_t mr_need = (std::max(C, L) + WAVE - 1) / WAVE;
    void (*merge_fn)(NndDev, uint32_t, uint32_t, int) =
        mr_need <= 2 ? nnd_merge_kernel<2> : mr_need <= 4 ? nnd_merge_kernel<4>
                                           : mr_need <= 8 ? nnd_merge_kernel<8> : nnd_merge_kernel<16>;
    NND_HIP(set_lds(merge_fn, merge_lds));
    NND_HIP(set_lds(nnd_rev_select_kernel, sel_lds));

    for (uint32_t it = 0; it < p->iters; ++it) {
        a.iter = it;
        float ms = 0.f;
        if (it > 0 || efanna) {
            if (it == 0) {  // EFANNA: the first join runs over random samples
                hipLaunchKernelGGL(nnd_efanna_seed_kernel, dim3(blocks), dim3(64 * WPB), WPB * 128 * 8, s, a);
                NND_HIP(hipGetLastError());
            }
            NND_HIP(hipEventRecord(ev[2], s));
            const uint32_t jb = (N + JOIN_WPB - 1) / JOIN_WPB;
            void *jargs[] = {&a};
            NND_HIP(hipLaunchKernel(reinterpret_cast<const void *>(join_fn), dim3(jb), dim3(64 * JOIN_WPB), jargs,
                                    join_lds, s));
            NND_HIP(hipEventRecord(ev[3], s));
            uint32_t up2_ = up2, hp2_ = hp2;
            int rehash = (int)(it + 1 < p->iters);
            void *margs[] = {&a, &up2_, &hp2_, &rehash};
            NND_HIP(hipLaunchKernel(reinterpret_cast<const void *>(merge_fn), dim3(blocks), dim3(64 * WPB), margs,
                                    merge_lds, s));
            NND_HIP(hipEventRecord(ev[4], s));
        } else {
            hipLaunchKernelGGL(nnd_m0_kernel, dim3(blocks), dim3(64 * WPB), 0, s, a);
            NND_HIP(hipGetLastError());
        }
        if (monitor) {  // after the round's merge: the pools update() would sort (outside the timed stages)
            const uint32_t mb = (monitor->n_ctrl + 3) / 4;
            hipLaunchKernelGGL(nnd_monitor_kernel, dim3(mb), dim3(256), 4 * monitor->truth_k * 4, s, a, *monitor, it);
            NND_HIP(hipGetLastError());
        }
        if (it + 1 < p->iters) {
            NND_HIP(hipEventRecord(ev[5], s));
            hipLaunchKernelGGL(fill_u32_kernel, dim3(1024), dim3(256), 0, s, a.rcnt_new, 0u, (size_t)N);
            hipLaunchKernelGGL(fill_u32_kernel, dim3(1024), dim3(256), 0, s, a.rcnt_old, 0u, (size_t)N);
            hipLaunchKernelGGL(nnd_sample_kernel, dim3(blocks), dim3(64 * WPB), 0, s, a);
            NND_HIP(hipGetLastError());
            NND_HIP(scan_u32(a.rcnt_new, N, a.roff_new, bsum, s));
            NND_HIP(scan_u32(a.rcnt_old, N, a.roff_old, bsum, s));
            hipLaunchKernelGGL(nnd_rev_scatter_kernel, dim3(blocks), dim3(64 * WPB), 0, s, a);
            NND_HIP(hipGetLastError());
            hipLaunchKernelGGL(nnd_rev_select_kernel, dim3(blocks), dim3(64 * WPB), sel_lds, s, a, rp2);
            NND_HIP(hipGetLastError());
            NND_HIP(hipEventRecord(ev[6], s));
        }
        NND_HIP(hipStreamSynchronize(s));
        if (it > 0 || efanna) {
            NND_HIP(hipEventElapsedTime(&ms, ev[2], ev[3]));
            t_join += ms;
            NND_HIP(hipEventElapsedTime(&ms, ev[3], ev[4]));
            t_merge += ms;
        }
        if (it + 1 < p->iters) {
            NND_HIP(hipEventElapsedTime(&ms, ev[5], ev[6]));
            t_sample += ms;
        }
    }
    NND_HIP(hipEventRecord(ev[2], s));
    switch (cpl) {
    case 1: NND_HIP(launch_final<1>(a, fp2, blocks, s)); break;
    case 2: NND_HIP(launch_final<2>(a, fp2, blocks, s)); break;
    case 4: NND_HIP(launch_final<4>(a, fp2, blocks, s)); break;
    case 8: NND_HIP(launch_final<8>(a, fp2, blocks, s)); break;
    default: NND_HIP(launch_final<16>(a, fp2, blocks, s)); break;
    }
    NND_HIP(hipEventRecord(ev[7], s));
    NND_HIP(hipStreamSynchronize(s));
    if (stats) {
        float ms = 0.f;
        std::vector<unsigned long long> cs((size_t)CTR_SLOTS * 4);
        NND_HIP(hipMemcpy(cs.data(), a.ctr, cs.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        unsigned long long c[4] = {0, 0, 0, 0};
        for (uint32_t i = 0; i < CTR_SLOTS; ++i)
            for (int j = 0; j < 4; ++j) c[j] += cs[(size_t)i * 4 + j];
        NND_HIP(hipEventElapsedTime(&ms, ev[2], ev[7]));
        stats->finalize_ms = ms;
        NND_HIP(hipEventElapsedTime(&ms, ev[0], ev[7]));
        stats->total_ms = ms;
        stats->init_ms = ms_init;
        stats->sample_ms = t_sample;
        stats->join_ms = t_join;
        stats->merge_ms = t_merge;
        stats->pairs = c[0];
        stats->proposals = c[1];
        stats->rows_gathered = c[2];
        stats->iters_run = p->iters;
        stats->cand_slots = C;
    }
    return WV_OK;
}

}  // namespace wv
