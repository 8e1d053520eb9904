// wv_exact.hip -- exact k nearest neighbours in FP64 (the recall oracle on
// the GPU; the only FP64 arithmetic in the library).
//
// Semantics follow the ground-truth generators of the reference
// (dataset/main.cpp:237-270, algo/knn/gen_knn.cpp:117-150): brute force over
// every base row, ascending distance, with two fixes: ties go to the smaller
// id, and the query's own row is skipped without leaving a hole (the
// reference's `continue` at gen_knn.cpp:136 leaves an unwritten slot).
//
// Each distance is sum_i ((double)q_i - (double)b_i)^2 accumulated in index
// order without fused multiply-add, the same expression and order as the
// CPU oracle, so the two agree bit for bit.
//
// Block = 256 threads = TQ queries x tiles of 256 base rows (one row per
// thread).  A candidate below its query's current k-th key is appended to
// that query's LDS buffer; after each tile one wavefront per query folds the
// buffer into the query's sorted top list with a bitonic sort.
#include "wv_device.hpp"
#include "wv_internal.hpp"

namespace wv {
namespace {

constexpr int EX_TQ = 8;
constexpr int EX_TB = 256;
constexpr int EX_KP = 128;  // list capacity, k <= EX_KP
constexpr int EX_SORT = 512;

struct XKey {
    uint64_t d;   // bits of a non-negative double: unsigned order == numeric order
    uint32_t id;
    uint32_t pad;
};
__device__ inline bool xless(const XKey &a, const XKey &b) {
    return a.d < b.d || (a.d == b.d && a.id < b.id);
}

__global__ void __launch_bounds__(256) exact_knn_kernel(ExactArgs a) {
    extern __shared__ __align__(16) uint8_t smem[];
    XKey *buf = reinterpret_cast<XKey *>(smem);                          // TQ x EX_SORT
    double *qs = reinterpret_cast<double *>(buf + EX_TQ * EX_SORT);      // TQ x d
    __shared__ uint32_t cnt[EX_TQ];
    __shared__ XKey thr[EX_TQ];

    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const uint32_t q0 = blockIdx.x * EX_TQ;
    const int nq = (int)min((uint32_t)EX_TQ, a.Q - q0);
    const uint32_t d = a.d;
    const XKey kmax = {~0ull, 0xFFFFFFFFu, 0};

    for (int i = t; i < EX_TQ * (int)d; i += blockDim.x) {
        int j = i / (int)d, c = i % (int)d;
        qs[i] = j < nq ? (double)a.queries[(size_t)(q0 + j) * a.ld_query + c] : 0.0;
    }
    for (int i = t; i < EX_TQ * EX_SORT; i += blockDim.x) buf[i] = kmax;
    if (t < EX_TQ) {
        cnt[t] = 0;
        thr[t] = kmax;
    }
    __syncthreads();

    for (uint32_t b0 = 0; b0 < a.N; b0 += EX_TB) {
        const uint32_t b = b0 + t;
        if (b < a.N) {
            double acc[EX_TQ];
#pragma unroll
            for (int j = 0; j < EX_TQ; ++j) acc[j] = 0.0;
            const float *row = a.base + (size_t)b * a.ld_base;
            for (uint32_t c = 0; c < d; ++c) {
                const double x = (double)row[c];
#pragma unroll
                for (int j = 0; j < EX_TQ; ++j) {
                    const double df = qs[j * d + c] - x;
                    acc[j] = acc[j] + df * df;
                }
            }
#pragma unroll
            for (int j = 0; j < EX_TQ; ++j) {
                if (j >= nq) continue;
                if (a.self_offset >= 0 && (int64_t)b == (int64_t)(q0 + j) + a.self_offset) continue;
                XKey k = {(uint64_t)__double_as_longlong(acc[j]), b, 0};
                if (xless(k, thr[j])) {
                    uint32_t pos = atomicAdd(&cnt[j], 1u);
                    buf[j * EX_SORT + EX_KP + pos] = k;
                }
            }
        }
        __syncthreads();
        for (int j = wid; j < nq; j += (int)(blockDim.x >> 6)) {
            if (cnt[j] == 0) continue;
            XKey *v = buf + j * EX_SORT;
            for (int k = 2; k <= EX_SORT; k <<= 1) {
                for (int jj = k >> 1; jj > 0; jj >>= 1) {
                    for (int i = lane; i < EX_SORT / 2; i += 64) {
                        int lo = ((i & ~(jj - 1)) << 1) | (i & (jj - 1));
                        int hi = lo | jj;
                        bool up = (lo & k) == 0;
                        XKey x = v[lo], y = v[hi];
                        if (xless(y, x) == up) {
                            v[lo] = y;
                            v[hi] = x;
                        }
                    }
                    wave_sync();
                }
            }
            for (int i = EX_KP + lane; i < EX_SORT; i += 64) v[i] = kmax;
            if (lane == 0) {
                cnt[j] = 0;
                thr[j] = v[a.k - 1];
            }
            wave_sync();
        }
        __syncthreads();
    }
    for (int j = wid; j < nq; j += (int)(blockDim.x >> 6)) {
        const XKey *v = buf + j * EX_SORT;
        for (uint32_t i = lane; i < a.k; i += 64) {
            a.out_ids[(size_t)(q0 + j) * a.k + i] = v[i].id;
            a.out_dist[(size_t)(q0 + j) * a.k + i] =
                v[i].id == 0xFFFFFFFFu ? -1.0 : __longlong_as_double((long long)v[i].d);
        }
    }
}

// Random row gathers in the search's distance phase, alone: each
// wavefront draws 64 uniformly random ids (a hash of wavefront, batch and
// lane, mod N) into LDS and scores their rows against a fixed query with
// the search's gather tile, `batches` times.  The dynamic LDS the launch
// asks for fixes how many wavefronts a CU holds.
template <int CPL, int G>
__global__ void __launch_bounds__(64) gather_rate_kernel(const float *base, uint32_t N, uint32_t dp, uint32_t nch,
                                                         uint32_t batches, uint64_t *out) {
    extern __shared__ __align__(16) uint8_t smem[];
    uint32_t *cid = reinterpret_cast<uint32_t *>(smem);
    uint64_t *ckey = reinterpret_cast<uint64_t *>(smem + WAVE * 4);
    const int lane = lane_id();
    const uint32_t w = blockIdx.x;
    float4 qv[CPL];
    load_query<CPL>(base, nch, qv);
    uint64_t acc = 0;
    for (uint32_t b = 0; b < batches; ++b) {
        cid[lane] = fmix32(w * 0x9E3779B1u + (b * WAVE + (uint32_t)lane) * 0x85EBCA6Bu) % N;
        wave_sync();
        rows_dist<CPL, G>(base, dp, nch, qv, cid, WAVE, ckey);
        wave_sync();
        acc ^= ckey[lane];
        wave_sync();
    }
    out[(size_t)w * WAVE + lane] = acc;
}

template <int CPL, int G>
hipError_t launch_gather_rate_t(const float *base, uint32_t N, uint32_t dp, uint32_t nch, uint32_t blocks,
                                uint32_t batches, uint32_t lds, uint64_t *out, hipStream_t s) {
    const void *fn = reinterpret_cast<const void *>(&gather_rate_kernel<CPL, G>);
    if (lds > 65536) {
        const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((gather_rate_kernel<CPL, G>), dim3(blocks), dim3(WAVE), lds, s, base, N, dp, nch, batches,
                       out);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_gather_rate(const float *base, uint32_t N, uint32_t dp, uint32_t blocks, uint32_t batches,
                              uint32_t lds, uint64_t *out, hipStream_t s) {
    const uint32_t nch = dp / 4;
    // the search's 64-row gather tile for each row length
    if (nch <= 16) return launch_gather_rate_t<1, 16>(base, N, dp, nch, blocks, batches, lds, out, s);
    if (nch <= 24) return launch_gather_rate_t<3, 8>(base, N, dp, nch, blocks, batches, lds, out, s);
    if (nch <= 32) return launch_gather_rate_t<2, 16>(base, N, dp, nch, blocks, batches, lds, out, s);
    if (nch <= 64) return launch_gather_rate_t<4, 8>(base, N, dp, nch, blocks, batches, lds, out, s);
    return hipErrorInvalidValue;
}

// Random 16-byte loads from an array of n16 16-byte vectors: each thread
// makes `iters` rounds of 4 independent loads at hashed indices (mod n16,
// so every access is inside the array).  What the GPU sustains for loads of
// random lines, the access pattern of the NN-Descent join's records, hash
// probes and slot updates.
__global__ void __launch_bounds__(256) line_rate_kernel(const uint4 *__restrict__ a, uint64_t n16, uint32_t iters,
                                                        uint32_t *out) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t h = fmix32(t * 0x9E3779B1u + 0x7F4A7C15u), acc = 0;
    for (uint32_t i = 0; i < iters; ++i) {
        uint4 v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            h = fmix32(h + 0x9E3779B9u);
            v[j] = a[(uint64_t)h % n16];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) acc ^= v[j].x ^ v[j].w;
    }
    out[t] = acc;
}

hipError_t launch_line_rate(const void *a, uint64_t bytes, uint32_t blocks, uint32_t iters, uint32_t *out,
                            hipStream_t s) {
    if (bytes < 16) return hipErrorInvalidValue;
    hipLaunchKernelGGL(line_rate_kernel, dim3(blocks), dim3(256), 0, s, static_cast<const uint4 *>(a), bytes / 16,
                       iters, out);
    return hipGetLastError();
}

hipError_t launch_exact_knn(const ExactArgs &a, hipStream_t s) {
    if (a.k == 0 || a.k > EX_KP) return hipErrorInvalidValue;
    const size_t lds = sizeof(XKey) * EX_TQ * EX_SORT + sizeof(double) * EX_TQ * a.d;
    if (lds > 160 * 1024 - 512) return hipErrorInvalidValue;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&exact_knn_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const uint32_t blocks = (a.Q + EX_TQ - 1) / EX_TQ;
    hipLaunchKernelGGL(exact_knn_kernel, dim3(blocks), dim3(256), lds, s, a);
    return hipGetLastError();
}

}  // namespace wv
