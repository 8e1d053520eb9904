// wv_capi.hip -- the C ABI (include/weavess_hip.h): contexts, device-resident
// indexes, argument checking and launch plumbing around the kernels.
#include "../../include/weavess_hip.h"
#include "wv_device.hpp"
#include "wv_internal.hpp"
#include "wv_mlp.hpp"
#include "wv_nnd.hpp"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <random>
#include <string>
#include <vector>

struct wv_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    void *scratch = nullptr;        // this device's copy of `scratch` (no-allocation search)
    hipError_t scratch_err = hipSuccess;
    std::mutex mu;
    std::string err;
    int cu_count = 0;
};

struct wv_index {
    wv_ctx *ctx = nullptr;
    uint32_t N = 0, d = 0, dp = 0;
    float *base = nullptr;      // N x dp
    uint32_t *graph = nullptr;  // N x deg_max
    float *gdist = nullptr;     // N x deg_max (NaN when unknown)
    uint32_t *deg = nullptr;    // N
    uint32_t deg_max = 0;
    wv_search_config cfg = {0, 0, 0, 0, 0, 0};
    uint32_t *queue = nullptr;  // work-queue counter of the persistent schedule (allocated on first use)
};

struct wv_search_graph;

namespace {

int fail(wv_ctx *c, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define WV_HIP(ctx, call)                                                                        \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail((ctx), e_ == hipErrorOutOfMemory ? WV_ENOMEM : WV_EHIP, "%s: %s (%s:%d)", \
                        #call, hipGetErrorString(e_), __FILE__, __LINE__);                       \
    } while (0)

uint32_t next_pow2(uint32_t v) {
    uint32_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Device buffer released at scope exit (used for per-call staging).
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 16); }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Copy `rows` host rows of d floats into a zero-padded device matrix of
// stride dp.
hipError_t upload_padded(float *dst, const float *src, uint32_t rows, uint32_t d, uint32_t dp,
                         hipStream_t s) {
    hipError_t e = hipSuccess;
    if (dp != d) e = hipMemsetAsync(dst, 0, (size_t)rows * dp * 4, s);
    if (e != hipSuccess) return e;
    if (rows == 0) return hipSuccess;
    return hipMemcpy2DAsync(dst, (size_t)dp * 4, src, (size_t)d * 4, (size_t)d * 4, rows,
                            hipMemcpyHostToDevice, s);
}

// Gather shape of the search kernels for rows of nch 4-element chunks: the
// chunks a lane holds, 16 lanes a row (1, 2, 4, 8, 16), or 8 lanes a row
// for 17-24 chunks (3).
int chunks_per_lane(uint32_t nch) {
    if (nch <= 16) return 1;
    if (nch <= 24) return 3;
    if (nch <= 32) return 2;
    if (nch <= 64) return 4;
    if (nch <= 128) return 8;
    if (nch <= 256) return 16;
    return 0;
}

struct SearchPlan {
    uint32_t P, H, B, wpb, groups, Lmax;
    bool persistent;
    uint32_t wpq;  // wavefronts per query (1, 2 or 4)
};

// Pools from this size up get the 256-key merge buffer by default.
constexpr uint32_t MERGE_BUFFER_MIN_L = 2560;

int plan_search(wv_index *ix, uint32_t Lmax, SearchPlan &pl, uint32_t C = 0) {
    pl.Lmax = Lmax;
    pl.P = std::max<uint32_t>(64, (Lmax + 63) / 64 * 64);
    // The visited table forgets under pressure (costing re-evaluations only);
    // LDS per wavefront sets how many queries are in flight, and that, not
    // the re-evaluations, bounds throughput on the benchmark graphs.
    // Default: large enough to stay exact for small pools; small for large
    // ones, where a bigger table buys < 1 % fewer evaluations but costs
    // resident queries.
    uint32_t H = ix->cfg.visited_slots;
    if (H == 0)
        H = Lmax <= 256 ? std::min<uint32_t>(2048, std::max<uint32_t>(1024, next_pow2(Lmax * 8)))
                        : (Lmax <= 512 ? 1024u : 256u);
    if (H < 64 || (H & (H - 1))) return fail(ix->ctx, WV_EINVAL, "visited_slots must be a power of two >= 64");
    pl.H = H;
    // merge buffer: pays off once a pool merge costs several 64-slot chunks
    const uint32_t mb = ix->cfg.merge_buffer;
    pl.B = mb == 0 ? (Lmax >= MERGE_BUFFER_MIN_L ? 256u : 0u) : (mb == 1 ? 0u : mb);
    uint32_t wpb = ix->cfg.waves_per_block ? ix->cfg.waves_per_block : 1;
    if (wpb != 1 && wpb != 2 && wpb != 4) return fail(ix->ctx, WV_EINVAL, "waves_per_block must be 1, 2 or 4");
    size_t lds = wv::search_wave_lds(pl.P, pl.H, pl.B, C);
    while (wpb > 1 && lds * wpb > 160 * 1024) wpb >>= 1;
    if (lds > 160 * 1024) return fail(ix->ctx, WV_ENOTSUP, "pool/visited table need %zu B of LDS", lds);
    pl.wpb = wpb;
    if (ix->cfg.schedule > 2) return fail(ix->ctx, WV_EINVAL, "schedule must be 0, 1 or 2");
    pl.wpq = ix->cfg.waves_per_query ? ix->cfg.waves_per_query : 1;
    if (pl.wpq != 1 && pl.wpq != 2 && pl.wpq != 4) return fail(ix->ctx, WV_EINVAL, "waves_per_query must be 1, 2 or 4");
    if (C) pl.wpq = 1;
    if (pl.wpq > 1 && wv::search_coop_lds(pl.P, pl.H, pl.B, pl.wpq) > 160 * 1024)
        return fail(ix->ctx, WV_ENOTSUP, "pool/visited table need %zu B of LDS",
                    wv::search_coop_lds(pl.P, pl.H, pl.B, pl.wpq));
    pl.persistent = ix->cfg.schedule == 2 && C == 0 && pl.wpq == 1;
    // gather tile: 4 row groups keep VGPRs (and so residency) low for d <= 256
    pl.groups = ix->cfg.row_groups ? ix->cfg.row_groups : (chunks_per_lane(ix->dp / 4) <= 4 ? 4u : 8u);
    return WV_OK;
}

// Core search on device buffers (caller holds the context lock).
struct SearchExtDev {
    const wv_term_model *model = nullptr;  // device copy
    float *trace = nullptr;
    uint32_t *trace_len = nullptr, *settle = nullptr;
    uint32_t cap = 0;
    uint64_t *full = nullptr;  // candidate collection (Q x C keys), or null
    uint32_t C = 0;
};

// Launch arguments of one search over device buffers (no allocation).
int prepare_search(wv_index *ix, const float *d_q, uint32_t Q, const uint32_t *d_entries, uint64_t seed,
                   uint32_t L, uint32_t k, const uint32_t *d_L, const uint32_t *d_k, uint32_t Lmax,
                   uint32_t k_stride, uint32_t *d_ids, float *d_dist, uint32_t *d_evals, uint32_t *d_hops,
                   const SearchExtDev &ext, wv::SearchArgs &a, SearchPlan &pl, int &cpl) {
    wv_ctx *c = ix->ctx;
    if (!ix->graph) return fail(c, WV_EINVAL, "index has no graph (build or set one first)");
    int rc = plan_search(ix, Lmax, pl, ext.C);
    if (rc) return rc;
    cpl = chunks_per_lane(ix->dp / 4);
    if (!cpl) return fail(c, WV_ENOTSUP, "dimension %u above 1024", ix->d);
    a.base = ix->base;
    a.queries = d_q;
    a.graph = ix->graph;
    a.deg = ix->deg;
    a.entries = d_entries;
    a.L_per_q = d_L;
    a.k_per_q = d_k;
    a.out_ids = d_ids;
    a.out_dist = d_dist;
    a.evals = d_evals;
    a.hops = d_hops;
    a.seed = seed;
    a.N = ix->N;
    a.dp = ix->dp;
    a.nch = ix->dp / 4;
    a.deg_max = ix->deg_max;
    a.Q = Q;
    a.L = L;
    a.k = k;
    a.k_stride = k_stride;
    a.P = pl.P;
    a.H = pl.H;
    a.B = pl.B;
    a.wave_lds = (uint32_t)wv::search_wave_lds(pl.P, pl.H, pl.B, ext.C);
    a.full = ext.full;
    a.C = ext.full ? ext.C : 0;
    a.model = ext.model;
    a.trace = ext.trace;
    a.trace_len = ext.trace_len;
    a.settle = ext.settle;
    a.trace_cap = ext.cap;
    a.queue = nullptr;
    a.order = nullptr;
    return WV_OK;
}

// The persistent schedule's work queue (the index's counter) and, given the
// per-query budgets, the order of its slots: largest L first (ties by query),
// so that the longest searches start first.  order_buf holds the order.
int attach_queue(wv_index *ix, const SearchPlan &pl, const uint32_t *h_L, uint32_t Q, DevBuf &order_buf,
                 wv::SearchArgs &a) {
    if (!pl.persistent) return WV_OK;
    wv_ctx *c = ix->ctx;
    if (!ix->queue) WV_HIP(c, hipMalloc(&ix->queue, 256));
    a.queue = ix->queue;
    if (h_L && Q > 1) {
        std::vector<uint32_t> ord(Q);
        std::iota(ord.begin(), ord.end(), 0u);
        std::stable_sort(ord.begin(), ord.end(), [h_L](uint32_t x, uint32_t y) { return h_L[x] > h_L[y]; });
        WV_HIP(c, order_buf.alloc((size_t)Q * 4));
        WV_HIP(c, hipMemcpyAsync(order_buf.p, ord.data(), (size_t)Q * 4, hipMemcpyHostToDevice, c->stream));
        WV_HIP(c, hipStreamSynchronize(c->stream));
        a.order = order_buf.as<uint32_t>();
    }
    return WV_OK;
}

// The plan's launch: W wavefronts a query when the plan asks for them and
// the search is plain, else one wavefront a query (grid or work queue).
hipError_t launch_plan(const wv::SearchArgs &a, const SearchPlan &pl, int cpl, wv_ctx *c) {
    if (pl.wpq > 1 && !a.model && !a.trace && !a.full && !a.queue)
        return wv::launch_greedy_search_coop(a, cpl, (int)pl.wpq, c->stream);
    return wv::launch_greedy_search(a, cpl, (int)pl.groups, (int)pl.wpb, c->stream, (uint32_t)c->cu_count);
}

int search_core(wv_index *ix, const float *d_q, uint32_t Q, const uint32_t *d_entries,
                uint64_t seed, uint32_t L, uint32_t k, const uint32_t *d_L, const uint32_t *d_k,
                uint32_t Lmax, uint32_t k_stride, uint32_t *d_ids, float *d_dist,
                uint32_t *d_evals, uint32_t *d_hops, wv_search_stats *stats,
                const SearchExtDev &ext = SearchExtDev(), const uint32_t *h_L = nullptr) {
    wv_ctx *c = ix->ctx;
    if (!ix->graph) return fail(c, WV_EINVAL, "index has no graph (build or set one first)");
    if (Q == 0) {
        if (stats) *stats = wv_search_stats{0.0, 0, 0};
        return WV_OK;
    }
    DevBuf ev_tmp, hop_tmp;
    if (stats && !d_evals) {
        WV_HIP(c, ev_tmp.alloc((size_t)Q * 4));
        d_evals = ev_tmp.as<uint32_t>();
    }
    if (stats && !d_hops) {
        WV_HIP(c, hop_tmp.alloc((size_t)Q * 4));
        d_hops = hop_tmp.as<uint32_t>();
    }
    wv::SearchArgs a{};
    SearchPlan pl;
    int cpl = 0;
    int rc = prepare_search(ix, d_q, Q, d_entries, seed, L, k, d_L, d_k, Lmax, k_stride, d_ids, d_dist, d_evals,
                            d_hops, ext, a, pl, cpl);
    if (rc) return rc;
    DevBuf order;
    if ((rc = attach_queue(ix, pl, h_L, Q, order, a))) return rc;
    WV_HIP(c, hipEventRecord(c->ev0, c->stream));
    WV_HIP(c, launch_plan(a, pl, cpl, c));
    WV_HIP(c, hipEventRecord(c->ev1, c->stream));
    WV_HIP(c, hipStreamSynchronize(c->stream));
    if (stats) {
        float ms = 0.f;
        WV_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        std::vector<uint32_t> ev(Q), hp(Q);
        WV_HIP(c, hipMemcpy(ev.data(), d_evals, (size_t)Q * 4, hipMemcpyDeviceToHost));
        WV_HIP(c, hipMemcpy(hp.data(), d_hops, (size_t)Q * 4, hipMemcpyDeviceToHost));
        stats->kernel_ms = ms;
        stats->dist_evals = 0;
        stats->hops = 0;
        for (uint32_t i = 0; i < Q; ++i) {
            stats->dist_evals += ev[i];
            stats->hops += hp[i];
        }
    }
    return WV_OK;
}

// Host-side argument checks shared by the search entry points.
int check_search_args(wv_ctx *c, uint32_t N, uint32_t Q, const void *entries, uint32_t L,
                      uint32_t k, const uint32_t *L_per_q, const uint32_t *k_per_q,
                      uint32_t k_stride, uint32_t &Lmax) {
    if ((L_per_q || k_per_q) && entries)
        return fail(c, WV_EINVAL, "explicit entries need one L for all queries");
    Lmax = 0;
    for (uint32_t q = 0; q < Q; ++q) {
        uint32_t Lq = L_per_q ? L_per_q[q] : L, kq = k_per_q ? k_per_q[q] : k;
        if (Lq == 0 || kq == 0)
            return fail(c, WV_EINVAL, "query %u: L and k must be positive", q);
        if (kq > Lq)  // src/builder.cpp:435-438
            return fail(c, WV_EINVAL, "query %u: search_L (%u) cannot be smaller than search_K (%u)", q, Lq, kq);
        if (kq > k_stride) return fail(c, WV_EINVAL, "query %u: k (%u) above k_stride (%u)", q, kq, k_stride);
        if (entries ? Lq > N : Lq >= N)
            return fail(c, WV_EINVAL, "query %u: L (%u) must be below N (%u) for random entries", q, Lq, N);
        Lmax = std::max(Lmax, Lq);
    }
    if (Lmax > 4096) return fail(c, WV_ENOTSUP, "L above 4096");
    return WV_OK;
}

// Explicit entry ids (host copy) must name base rows.
int check_entries(wv_ctx *c, const uint32_t *entries, size_t n, uint32_t N) {
    for (size_t i = 0; i < n; ++i)
        if (entries[i] >= N) return fail(c, WV_EINVAL, "entry %zu is %u, not below N=%u", i, entries[i], N);
    return WV_OK;
}

// Address of a member of a device-resident model (no host dereference).
const float *model_field(const void *dev_model, size_t offset) {
    return reinterpret_cast<const float *>(static_cast<const char *>(dev_model) + offset);
}

}  // namespace

struct wv_search_graph {
    wv_index *ix = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void *model = nullptr;  // device copy of the stop model, or null
    DevBuf order;           // persistent schedule: the slots' query order, or empty
};

// ------------------------------------------------------------ no allocation
// Scratch of wv_index_search_dev_noalloc: the per-query evaluation and hop
// counts a caller does not take, then the two totals of the statistics.  It
// is file-scope device memory: each GPU the library runs on holds its own
// copy, at its own address, until the process ends.  Host code never passes
// the symbol itself to a kernel or a copy (on the host it only names the
// array); it uses the address hipGetSymbolAddress returned for the device.
constexpr size_t NOALLOC_COUNTS = 2 * (size_t)WV_NOALLOC_MAX_Q * 4;
static __device__ __align__(256) uint8_t scratch[NOALLOC_COUNTS + 256];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch[MAX_DEVICES];          // per device: the address of `scratch`
std::mutex g_scratch_call_mu[MAX_DEVICES];  // calls on one GPU share its array

// Once per device, with that device current (wv_ctx_create).
hipError_t scratch_address(int device, void **out) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    *out = nullptr;
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = p;
    }
    *out = g_scratch[device];
    return hipSuccess;
}

__global__ void totals_clear_kernel(unsigned long long *tot) {
    if (threadIdx.x < 2) tot[threadIdx.x] = 0;
}

// tot[0] += sum of ev[0..Q), tot[1] += sum of hp[0..Q)
__global__ void __launch_bounds__(256) totals_kernel(const uint32_t *ev, const uint32_t *hp, uint32_t Q,
                                                     unsigned long long *tot) {
    uint64_t se = 0, sh = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < Q; i += gridDim.x * blockDim.x) {
        se += ev[i];
        sh += hp[i];
    }
    for (int o = wv::WAVE / 2; o > 0; o >>= 1) {
        se += wv::shfl_xor64(se, o);
        sh += wv::shfl_xor64(sh, o);
    }
    if ((threadIdx.x & (wv::WAVE - 1)) == 0) {
        atomicAdd(&tot[0], (unsigned long long)se);
        atomicAdd(&tot[1], (unsigned long long)sh);
    }
}

}  // namespace

extern "C" {

int wv_version(void) { return 1; }

int wv_ctx_create(int device, wv_ctx **out) {
    if (!out) return WV_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return WV_EHIP;
    if (device < 0 || device >= n) return WV_EINVAL;
    wv_ctx *c = new (std::nothrow) wv_ctx;
    if (!c) return WV_ENOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return WV_EHIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cu_count = prop.multiProcessorCount;
    // the no-allocation search's scratch: its address on this device, found
    // once per device while it is current (a failure is reported by that
    // entry point only)
    c->scratch_err = scratch_address(device, &c->scratch);
    *out = c;
    return WV_OK;
}

int wv_ctx_destroy(wv_ctx *c) {
    if (!c) return WV_EINVAL;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return WV_OK;
}

const char *wv_last_error(const wv_ctx *c) { return c ? c->err.c_str() : "null context"; }

int wv_index_create(wv_ctx *c, const float *base, uint32_t N, uint32_t d, wv_index **out) {
    if (!c || !out) return WV_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    *out = nullptr;
    if (!base) return fail(c, WV_EINVAL, "base is null");
    if (N < 2 || d == 0) return fail(c, WV_EINVAL, "need N >= 2 and d >= 1 (N=%u d=%u)", N, d);
    // the search keeps a flag in bit 31 of each pooled id
    if (N > 0x7FFFFFFFu) return fail(c, WV_ENOTSUP, "N above 2^31 - 1");
    if (d > 1024) return fail(c, WV_ENOTSUP, "dimension %u above 1024", d);
    WV_HIP(c, hipSetDevice(c->device));
    wv_index *ix = new (std::nothrow) wv_index;
    if (!ix) return fail(c, WV_ENOMEM, "host allocation");
    ix->ctx = c;
    ix->N = N;
    ix->d = d;
    ix->dp = (d + 3) & ~3u;
    hipError_t e = hipMalloc(&ix->base, (size_t)N * ix->dp * 4);
    if (e == hipSuccess) e = upload_padded(ix->base, base, N, d, ix->dp, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        if (ix->base) (void)hipFree(ix->base);
        delete ix;
        return fail(c, e == hipErrorOutOfMemory ? WV_ENOMEM : WV_EHIP, "index upload: %s", hipGetErrorString(e));
    }
    *out = ix;
    return WV_OK;
}

int wv_index_destroy(wv_index *ix) {
    if (!ix) return WV_EINVAL;
    std::lock_guard<std::mutex> g(ix->ctx->mu);
    (void)hipSetDevice(ix->ctx->device);
    if (ix->queue) (void)hipFree(ix->queue);
    if (ix->base) (void)hipFree(ix->base);
    if (ix->graph) (void)hipFree(ix->graph);
    if (ix->gdist) (void)hipFree(ix->gdist);
    if (ix->deg) (void)hipFree(ix->deg);
    delete ix;
    return WV_OK;
}

}  // extern "C"

namespace wv {
// Replace the index's graph storage with freshly allocated N x deg_max
// buffers (contents undefined).
int index_alloc_graph(wv_index *ix, uint32_t deg_max) {
    wv_ctx *c = ix->ctx;
    if (ix->graph) (void)hipFree(ix->graph);
    if (ix->gdist) (void)hipFree(ix->gdist);
    if (ix->deg) (void)hipFree(ix->deg);
    ix->graph = nullptr;
    ix->gdist = nullptr;
    ix->deg = nullptr;
    ix->deg_max = 0;
    WV_HIP(c, hipMalloc(&ix->graph, (size_t)ix->N * std::max<uint32_t>(deg_max, 1) * 4));
    WV_HIP(c, hipMalloc(&ix->gdist, (size_t)ix->N * std::max<uint32_t>(deg_max, 1) * 4));
    WV_HIP(c, hipMalloc(&ix->deg, (size_t)ix->N * 4));
    ix->deg_max = deg_max;
    return WV_OK;
}
IndexView index_view(wv_index *ix) {
    IndexView v;
    v.base = ix->base;
    v.N = ix->N;
    v.d = ix->d;
    v.dp = ix->dp;
    v.stream = ix->ctx->stream;
    return v;
}
int index_set_error(wv_index *ix, int code, const char *msg) { return fail(ix->ctx, code, "%s", msg); }

int search_collect(wv_index *ix, const uint32_t *graph, const uint32_t *deg, uint32_t deg_max, const float *d_q,
                   uint32_t Q, const uint32_t *d_entries, uint32_t L, uint32_t C, uint64_t *d_full,
                   uint32_t *d_evals) {
    wv_ctx *c = ix->ctx;
    if (Q == 0) return WV_OK;
    uint32_t *g_save = ix->graph, *d_save = ix->deg, m_save = ix->deg_max, h_save = ix->cfg.visited_slots;
    ix->graph = const_cast<uint32_t *>(graph);
    ix->deg = const_cast<uint32_t *>(deg);
    ix->deg_max = deg_max;
    // The collected candidates do not depend on the visited table; at
    // L <= 256 a 1024-slot table leaves more queries in flight than the
    // 2048 slots that keep plain search's evaluation counts exact (NSG at
    // L=150 on 1M x 128: 0.45 s of searching instead of 0.65 s).  At
    // L <= 80 a 512-slot table evicts little (Vamana at L=70: 0.33 s instead
    // of 0.40 s, the same evaluations); at L=150 it re-evaluates 15 % more.
    if (h_save == 0 && L <= 256) ix->cfg.visited_slots = L <= 80 ? 512 : 1024;
    struct Restore {
        wv_index *ix;
        uint32_t *g, *d, m, h;
        ~Restore() {
            ix->graph = g;
            ix->deg = d;
            ix->deg_max = m;
            ix->cfg.visited_slots = h;
        }
    } restore{ix, g_save, d_save, m_save, h_save};
    DevBuf ids;
    WV_HIP(c, ids.alloc((size_t)Q * 4));
    SearchExtDev x;
    x.full = d_full;
    x.C = C;
    wv::SearchArgs a{};
    SearchPlan pl;
    int cpl = 0;
    int rc = prepare_search(ix, d_q, Q, d_entries, 0, L, 1, nullptr, nullptr, L, 1, ids.as<uint32_t>(), nullptr,
                            d_evals, nullptr, x, a, pl, cpl);
    if (rc) return rc;
    WV_HIP(c, wv::launch_greedy_search(a, cpl, (int)pl.groups, (int)pl.wpb, c->stream));
    WV_HIP(c, hipStreamSynchronize(c->stream));
    return WV_OK;
}

int index_alloc_graph_dev(wv_index *ix, uint32_t W, uint32_t **ids, float **dist, uint32_t **deg) {
    int rc = index_alloc_graph(ix, W);
    if (rc) return rc;
    *ids = ix->graph;
    *dist = ix->gdist;
    *deg = ix->deg;
    return WV_OK;
}

int index_upload_graph(wv_index *ix, const uint32_t *ids, const float *dist, const uint32_t *deg, uint32_t W) {
    wv_ctx *c = ix->ctx;
    int rc = index_alloc_graph(ix, W);
    if (rc) return rc;
    WV_HIP(c, hipMemcpy(ix->graph, ids, (size_t)ix->N * W * 4, hipMemcpyHostToDevice));
    WV_HIP(c, hipMemcpy(ix->gdist, dist, (size_t)ix->N * W * 4, hipMemcpyHostToDevice));
    WV_HIP(c, hipMemcpy(ix->deg, deg, (size_t)ix->N * 4, hipMemcpyHostToDevice));
    return WV_OK;
}
}  // namespace wv

extern "C" {

int wv_index_set_graph(wv_index *ix, const uint32_t *ids, const uint32_t *deg, uint32_t deg_max) {
    if (!ix) return WV_EINVAL;
    wv_ctx *c = ix->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    if (!ids || !deg) return fail(c, WV_EINVAL, "graph ids/deg are null");
    if (deg_max == 0) return fail(c, WV_EINVAL, "deg_max must be positive");
    WV_HIP(c, hipSetDevice(c->device));
    int rc = wv::index_alloc_graph(ix, deg_max);
    if (rc) return rc;
    WV_HIP(c, hipMemcpyAsync(ix->graph, ids, (size_t)ix->N * deg_max * 4, hipMemcpyHostToDevice, c->stream));
    WV_HIP(c, hipMemcpyAsync(ix->deg, deg, (size_t)ix->N * 4, hipMemcpyHostToDevice, c->stream));
    WV_HIP(c, hipMemsetAsync(ix->gdist, 0xFF, (size_t)ix->N * deg_max * 4, c->stream));
    WV_HIP(c, hipStreamSynchronize(c->stream));
    return WV_OK;
}

int wv_index_get_graph(wv_index *ix, uint32_t *ids, float *dist, uint32_t *deg, uint32_t *deg_max) {
    if (!ix) return WV_EINVAL;
    wv_ctx *c = ix->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    if (!ix->graph) return fail(c, WV_EINVAL, "index has no graph");
    if (deg_max) *deg_max = ix->deg_max;
    WV_HIP(c, hipSetDevice(c->device));
    if (ids) WV_HIP(c, hipMemcpy(ids, ix->graph, (size_t)ix->N * ix->deg_max * 4, hipMemcpyDeviceToHost));
    if (dist) WV_HIP(c, hipMemcpy(dist, ix->gdist, (size_t)ix->N * ix->deg_max * 4, hipMemcpyDeviceToHost));
    if (deg) WV_HIP(c, hipMemcpy(deg, ix->deg, (size_t)ix->N * 4, hipMemcpyDeviceToHost));
    return WV_OK;
}

int wv_index_set_search_config(wv_index *ix, const wv_search_config *cfg) {
    if (!ix || !cfg) return WV_EINVAL;
    std::lock_guard<std::mutex> g(ix->ctx->mu);
    if (cfg->visited_slots && (cfg->visited_slots < 64 || (cfg->visited_slots & (cfg->visited_slots - 1))))
        return fail(ix->ctx, WV_EINVAL, "visited_slots must be 0 or a power of two >= 64");
    if (cfg->waves_per_block > 4 || cfg->waves_per_block == 3)
        return fail(ix->ctx, WV_EINVAL, "waves_per_block must be 0, 1, 2 or 4");
    if (cfg->merge_buffer > 1 && (cfg->merge_buffer % 64 || cfg->merge_buffer > 256))
        return fail(ix->ctx, WV_EINVAL, "merge_buffer must be 0, 1, 64, 128, 192 or 256");
    if (cfg->schedule > 2) return fail(ix->ctx, WV_EINVAL, "schedule must be 0, 1 or 2");
    if (cfg->waves_per_query > 4 || cfg->waves_per_query == 3)
        return fail(ix->ctx, WV_EINVAL, "waves_per_query must be 0, 1, 2 or 4");
    ix->cfg = *cfg;
    return WV_OK;
}

int wv_index_search(wv_index *ix, const float *queries, uint32_t Q, const uint32_t *entries,
                    uint64_t seed, uint32_t L, uint32_t k, const uint32_t *L_per_q,
                    const uint32_t *k_per_q, uint32_t k_stride, uint32_t *out_ids, float *out_dist,
                    uint32_t *evals, uint32_t *hops, wv_search_stats *stats) {
    if (!ix) return WV_EINVAL;
    wv_ctx *c = ix->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    if (!queries || !out_ids) return fail(c, WV_EINVAL, "queries/out_ids are null");
    uint32_t Lmax = 0;
    int rc = check_search_args(c, ix->N, Q, entries, L, k, L_per_q, k_per_q, k_stride, Lmax);
    if (rc) return rc;
    if (entries && (rc = check_entries(c, entries, (size_t)Q * L, ix->N))) return rc;
    WV_HIP(c, hipSetDevice(c->device));
    if (Q == 0) {
        if (stats) *stats = wv_search_stats{0.0, 0, 0};
        return WV_OK;
    }
    DevBuf dq, de, dl, dk, di, dd, dev, dh;
    WV_HIP(c, dq.alloc((size_t)Q * ix->dp * 4));
    WV_HIP(c, upload_padded(dq.as<float>(), queries, Q, ix->d, ix->dp, c->stream));
    if (entries) {
        WV_HIP(c, de.alloc((size_t)Q * L * 4));
        WV_HIP(c, hipMemcpyAsync(de.p, entries, (size_t)Q * L * 4, hipMemcpyHostToDevice, c->stream));
    }
    if (L_per_q) {
        WV_HIP(c, dl.alloc((size_t)Q * 4));
        WV_HIP(c, hipMemcpyAsync(dl.p, L_per_q, (size_t)Q * 4, hipMemcpyHostToDevice, c->stream));
    }
    if (k_per_q) {
        WV_HIP(c, dk.alloc((size_t)Q * 4));
        WV_HIP(c, hipMemcpyAsync(dk.p, k_per_q, (size_t)Q * 4, hipMemcpyHostToDevice, c->stream));
    }
    WV_HIP(c, di.alloc((size_t)Q * k_stride * 4));
    if (out_dist) WV_HIP(c, dd.alloc((size_t)Q * k_stride * 4));
    if (evals) WV_HIP(c, dev.alloc((size_t)Q * 4));
    if (hops) WV_HIP(c, dh.alloc((size_t)Q * 4));
    rc = search_core(ix, dq.as<float>(), Q, de.as<uint32_t>(), seed, L, k, dl.as<uint32_t>(),
                     dk.as<uint32_t>(), Lmax, k_stride, di.as<uint32_t>(), dd.as<float>(),
                     dev.as<uint32_t>(), dh.as<uint32_t>(), stats, SearchExtDev(), L_per_q);
    if (rc) return rc;
    WV_HIP(c, hipMemcpy(out_ids, di.p, (size_t)Q * k_stride * 4, hipMemcpyDeviceToHost));
    if (out_dist) WV_HIP(c, hipMemcpy(out_dist, dd.p, (size_t)Q * k_stride * 4, hipMemcpyDeviceToHost));
    if (evals) WV_HIP(c, hipMemcpy(evals, dev.p, (size_t)Q * 4, hipMemcpyDeviceToHost));
    if (hops) WV_HIP(c, hipMemcpy(hops, dh.p, (size_t)Q * 4, hipMemcpyDeviceToHost));
    return WV_OK;
}

int wv_index_search_dev(wv_index *ix, const float *d_queries, uint32_t Q, const uint32_t *d_entries,
                        uint64_t seed, uint32_t L, uint32_t k, const uint32_t *d_L_per_q,
                        const uint32_t *d_k_per_q, uint32_t k_stride, uint32_t *d_out_ids,
                        float *d_out_dist, uint32_t *d_evals, uint32_t *d_hops,
                        wv_search_stats *stats) {
    if (!ix) return WV_EINVAL;
    wv_ctx *c = ix->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    if (!d_queries || !d_out_ids) return fail(c, WV_EINVAL, "queries/out_ids are null");
    WV_HIP(c, hipSetDevice(c->device));
    // per-query L/k are checked on the host before any launch
    std::vector<uint32_t> hl, hk;
    if (d_L_per_q) {
        hl.resize(Q);
        WV_HIP(c, hipMemcpy(hl.data(), d_L_per_q, (size_t)Q * 4, hipMemcpyDeviceToHost));
    }
    if (d_k_per_q) {
        hk.resize(Q);
        WV_HIP(c, hipMemcpy(hk.data(), d_k_per_q, (size_t)Q * 4, hipMemcpyDeviceToHost));
    }
    uint32_t Lmax = 0;
    int rc = check_search_args(c, ix->N, Q, d_entries, L, k, d_L_per_q ? hl.data() : nullptr,
                               d_k_per_q ? hk.data() : nullptr, k_stride, Lmax);
    if (rc) return rc;
    if (d_entries && Q) {
        std::vector<uint32_t> he((size_t)Q * L);
        WV_HIP(c, hipMemcpy(he.data(), d_entries, he.size() * 4, hipMemcpyDeviceToHost));
        if ((rc = check_entries(c, he.data(), he.size(), ix->N))) return rc;
    }
    return search_core(ix, d_queries, Q, d_entries, seed, L, k, d_L_per_q, d_k_per_q, Lmax,
                       k_stride, d_out_ids, d_out_dist, d_evals, d_hops, stats, SearchExtDev(),
                       d_L_per_q ? hl.data() : nullptr);
}

int wv_index_search_dev_noalloc(wv_index *ix, const float *d_queries, uint32_t Q, uint64_t seed, uint32_t L,
                                uint32_t k, uint32_t k_stride, uint32_t *d_out_ids, float *d_out_dist,
                                uint32_t *d_evals, uint32_t *d_hops, wv_search_stats *stats) {
    if (!ix) return WV_EINVAL;
    wv_ctx *c = ix->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    if (!d_queries || !d_out_ids) return fail(c, WV_EINVAL, "queries/out_ids are null");
    if (Q > WV_NOALLOC_MAX_Q)
        return fail(c, WV_EINVAL, "no-allocation search: %u queries, above its maximum of %u a call", Q,
                    (unsigned)WV_NOALLOC_MAX_Q);
    uint32_t Lmax = 0;
    int rc = check_search_args(c, ix->N, Q, nullptr, L, k, nullptr, nullptr, k_stride, Lmax);
    if (rc) return rc;
    if (!ix->graph) return fail(c, WV_EINVAL, "index has no graph (build or set one first)");
    if (!c->scratch)
        return fail(c, WV_EHIP, "no-allocation search: scratch address unavailable on device %d (%s)", c->device,
                    hipGetErrorString(c->scratch_err));
    WV_HIP(c, hipSetDevice(c->device));
    if (Q == 0) {
        if (stats) *stats = wv_search_stats{0.0, 0, 0};
        return WV_OK;
    }
    std::lock_guard<std::mutex> g2(g_scratch_call_mu[c->device]);
    uint32_t *s_ev = static_cast<uint32_t *>(c->scratch);
    uint32_t *s_hp = s_ev + WV_NOALLOC_MAX_Q;
    auto *tot = reinterpret_cast<unsigned long long *>(static_cast<uint8_t *>(c->scratch) + NOALLOC_COUNTS);
    uint32_t *ev = d_evals ? d_evals : (stats ? s_ev : nullptr);
    uint32_t *hp = d_hops ? d_hops : (stats ? s_hp : nullptr);
    wv::SearchArgs a{};
    SearchPlan pl;
    int cpl = 0;
    rc = prepare_search(ix, d_queries, Q, nullptr, seed, L, k, nullptr, nullptr, Lmax, k_stride, d_out_ids,
                        d_out_dist, ev, hp, SearchExtDev(), a, pl, cpl);
    if (rc) return rc;
    // the persistent schedule's work-queue counter is in the array too (the
    // launch zeroes it)
    if (pl.persistent) a.queue = reinterpret_cast<uint32_t *>(tot + 2);
    // the totals start at zero on every call (the array keeps the last call's)
    if (stats) {
        hipLaunchKernelGGL(totals_clear_kernel, dim3(1), dim3(64), 0, c->stream, tot);
        WV_HIP(c, hipGetLastError());
    }
    WV_HIP(c, hipEventRecord(c->ev0, c->stream));
    WV_HIP(c, launch_plan(a, pl, cpl, c));
    WV_HIP(c, hipEventRecord(c->ev1, c->stream));
    unsigned long long h_tot[2] = {0, 0};
    if (stats) {
        const uint32_t blocks = std::min<uint32_t>(1024, (Q + 255) / 256);
        hipLaunchKernelGGL(totals_kernel, dim3(blocks), dim3(256), 0, c->stream, ev, hp, Q, tot);
        WV_HIP(c, hipGetLastError());
        WV_HIP(c, hipMemcpyAsync(h_tot, tot, sizeof h_tot, hipMemcpyDeviceToHost, c->stream));
    }
    WV_HIP(c, hipStreamSynchronize(c->stream));
    if (stats) {
        float ms = 0.f;
        WV_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        stats->kernel_ms = ms;
        stats->dist_evals = h_tot[0];
        stats->hops = h_tot[1];
    }
    return WV_OK;
}

int wv_index_search_ext(wv_index *ix, const float *queries, uint32_t Q, const uint32_t *entries,
                        uint64_t seed, uint32_t L, uint32_t k, const uint32_t *L_per_q,
                        const uint32_t *k_per_q, uint32_t k_stride, uint32_t *out_ids,
                        float *out_dist, uint32_t *evals, uint32_t *hops,
                        const wv_search_ext *ext, wv_search_stats *stats) {
    if (!ix) return WV_EINVAL;
    wv_ctx *c = ix->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    if (!queries || !out_ids) return fail(c, WV_EINVAL, "queries/out_ids are null");
    if (ext && ext->trace && (!ext->trace_len || ext->trace_cap == 0))
        return fail(c, WV_EINVAL, "trace needs trace_len and trace_cap > 0");
    uint32_t Lmax = 0;
    int rc = check_search_args(c, ix->N, Q, entries, L, k, L_per_q, k_per_q, k_stride, Lmax);
    if (rc) return rc;
    if (entries && (rc = check_entries(c, entries, (size_t)Q * L, ix->N))) return rc;
    WV_HIP(c, hipSetDevice(c->device));
    if (Q == 0) {
        if (stats) *stats = wv_search_stats{0.0, 0, 0};
        return WV_OK;
    }
    DevBuf dq, de, dl, dk, di, dd, dev, dh, dm, dtr, dtl, dse;
    WV_HIP(c, dq.alloc((size_t)Q * ix->dp * 4));
    WV_HIP(c, upload_padded(dq.as<float>(), queries, Q, ix->d, ix->dp, c->stream));
    if (entries) {
        WV_HIP(c, de.alloc((size_t)Q * L * 4));
        WV_HIP(c, hipMemcpyAsync(de.p, entries, (size_t)Q * L * 4, hipMemcpyHostToDevice, c->stream));
    }
    if (L_per_q) {
        WV_HIP(c, dl.alloc((size_t)Q * 4));
        WV_HIP(c, hipMemcpyAsync(dl.p, L_per_q, (size_t)Q * 4, hipMemcpyHostToDevice, c->stream));
    }
    if (k_per_q) {
        WV_HIP(c, dk.alloc((size_t)Q * 4));
        WV_HIP(c, hipMemcpyAsync(dk.p, k_per_q, (size_t)Q * 4, hipMemcpyHostToDevice, c->stream));
    }
    WV_HIP(c, di.alloc((size_t)Q * k_stride * 4));
    if (out_dist) WV_HIP(c, dd.alloc((size_t)Q * k_stride * 4));
    if (evals) WV_HIP(c, dev.alloc((size_t)Q * 4));
    if (hops) WV_HIP(c, dh.alloc((size_t)Q * 4));
    SearchExtDev x;
    if (ext && ext->model) {
        WV_HIP(c, dm.alloc(sizeof(wv_term_model)));
        WV_HIP(c, hipMemcpyAsync(dm.p, ext->model, sizeof(wv_term_model), hipMemcpyHostToDevice, c->stream));
        x.model = dm.as<wv_term_model>();
    }
    if (ext && ext->trace) {
        x.cap = ext->trace_cap;
        WV_HIP(c, dtr.alloc((size_t)Q * x.cap * 8 * 4));
        WV_HIP(c, dtl.alloc((size_t)Q * 4));
        x.trace = dtr.as<float>();
        x.trace_len = dtl.as<uint32_t>();
    }
    if (ext && ext->settle_hop) {
        WV_HIP(c, dse.alloc((size_t)Q * 4));
        x.settle = dse.as<uint32_t>();
    }
    rc = search_core(ix, dq.as<float>(), Q, de.as<uint32_t>(), seed, L, k, dl.as<uint32_t>(),
                     dk.as<uint32_t>(), Lmax, k_stride, di.as<uint32_t>(), dd.as<float>(),
                     dev.as<uint32_t>(), dh.as<uint32_t>(), stats, x, L_per_q);
    if (rc) return rc;
    WV_HIP(c, hipMemcpy(out_ids, di.p, (size_t)Q * k_stride * 4, hipMemcpyDeviceToHost));
    if (out_dist) WV_HIP(c, hipMemcpy(out_dist, dd.p, (size_t)Q * k_stride * 4, hipMemcpyDeviceToHost));
    if (evals) WV_HIP(c, hipMemcpy(evals, dev.p, (size_t)Q * 4, hipMemcpyDeviceToHost));
    if (hops) WV_HIP(c, hipMemcpy(hops, dh.p, (size_t)Q * 4, hipMemcpyDeviceToHost));
    if (x.trace) {
        WV_HIP(c, hipMemcpy(ext->trace, dtr.p, (size_t)Q * x.cap * 8 * 4, hipMemcpyDeviceToHost));
        WV_HIP(c, hipMemcpy(ext->trace_len, dtl.p, (size_t)Q * 4, hipMemcpyDeviceToHost));
    }
    if (x.settle) WV_HIP(c, hipMemcpy(ext->settle_hop, dse.p, (size_t)Q * 4, hipMemcpyDeviceToHost));
    return WV_OK;
}

int wv_index_search_ext_dev(wv_index *ix, const float *d_queries, uint32_t Q, const uint32_t *d_entries,
                            uint64_t seed, uint32_t L, uint32_t k, const uint32_t *d_L_per_q,
                            const uint32_t *d_k_per_q, uint32_t k_stride, uint32_t *d_out_ids,
                            float *d_out_dist, uint32_t *d_evals, uint32_t *d_hops,
                            const wv_search_ext *ext, wv_search_stats *stats) {
    if (!ix) return WV_EINVAL;
    wv_ctx *c = ix->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    if (!d_queries || !d_out_ids) return fail(c, WV_EINVAL, "queries/out_ids are null");
    if (ext && ext->trace && (!ext->trace_len || ext->trace_cap == 0))
        return fail(c, WV_EINVAL, "trace needs trace_len and trace_cap > 0");
    if (ext && ext->trace && (reinterpret_cast<uintptr_t>(ext->trace) & 15))
        return fail(c, WV_EINVAL, "trace must be 16-byte aligned");
    WV_HIP(c, hipSetDevice(c->device));
    std::vector<uint32_t> hl, hk;
    if (d_L_per_q) {
        hl.resize(Q);
        WV_HIP(c, hipMemcpy(hl.data(), d_L_per_q, (size_t)Q * 4, hipMemcpyDeviceToHost));
    }
    if (d_k_per_q) {
        hk.resize(Q);
        WV_HIP(c, hipMemcpy(hk.data(), d_k_per_q, (size_t)Q * 4, hipMemcpyDeviceToHost));
    }
    uint32_t Lmax = 0;
    int rc = check_search_args(c, ix->N, Q, d_entries, L, k, d_L_per_q ? hl.data() : nullptr,
                               d_k_per_q ? hk.data() : nullptr, k_stride, Lmax);
    if (rc) return rc;
    if (d_entries && Q) {
        std::vector<uint32_t> he((size_t)Q * L);
        WV_HIP(c, hipMemcpy(he.data(), d_entries, he.size() * 4, hipMemcpyDeviceToHost));
        if ((rc = check_entries(c, he.data(), he.size(), ix->N))) return rc;
    }
    DevBuf dm;
    SearchExtDev x;
    if (ext && ext->model) {
        WV_HIP(c, dm.alloc(sizeof(wv_term_model)));
        WV_HIP(c, hipMemcpy(dm.p, ext->model, sizeof(wv_term_model), hipMemcpyHostToDevice));
        x.model = dm.as<wv_term_model>();
    }
    if (ext) {
        x.trace = ext->trace;
        x.trace_len = ext->trace ? ext->trace_len : nullptr;
        x.cap = ext->trace ? ext->trace_cap : 0;
        x.settle = ext->settle_hop;
    }
    return search_core(ix, d_queries, Q, d_entries, seed, L, k, d_L_per_q, d_k_per_q, Lmax, k_stride,
                       d_out_ids, d_out_dist, d_evals, d_hops, stats, x, d_L_per_q ? hl.data() : nullptr);
}

}  // extern "C"

namespace {

// The autotuner over one L and k, or over per-query budgets (d_L/d_k, with
// their host copies hl/hk, already checked): coordinate descent from the
// default plan, each candidate the best of 3 timed launches after a warm-up.
int autotune_core(wv_index *ix, const float *d_queries, uint32_t Q, uint64_t seed, uint32_t L, uint32_t k,
                  const uint32_t *d_L, const uint32_t *d_k, const uint32_t *hl, uint32_t k_stride, uint32_t Lmax,
                  const wv_search_ext *ext, wv_search_config *best, double *best_ms) {
    wv_ctx *c = ix->ctx;
    WV_HIP(c, hipSetDevice(c->device));
    DevBuf dm, di;
    SearchExtDev x;
    if (ext && ext->model) {
        WV_HIP(c, dm.alloc(sizeof(wv_term_model)));
        WV_HIP(c, hipMemcpy(dm.p, ext->model, sizeof(wv_term_model), hipMemcpyHostToDevice));
        x.model = dm.as<wv_term_model>();
    }
    WV_HIP(c, di.alloc((size_t)Q * k_stride * 4));
    const wv_search_config saved = ix->cfg;
    // best of 3 timed launches after a warm-up; planning errors reject the
    // candidate (rc EINVAL/ENOTSUP), launch errors end the tuning
    auto time_cfg = [&](const wv_search_config &cfg, double &ms) -> int {
        ix->cfg = cfg;
        wv_search_stats st;
        ms = 1e30;
        for (int r = 0; r < 4; ++r) {
            const int e = search_core(ix, d_queries, Q, nullptr, seed, L, k, d_L, d_k, Lmax, k_stride,
                                      di.as<uint32_t>(), nullptr, nullptr, nullptr, &st, x, hl);
            if (e) return e;
            if (r > 0) ms = std::min(ms, st.kernel_ms);
        }
        return WV_OK;
    };
    wv_search_config cur = {0, 0, 0, 0, 0, 0};
    double cur_ms = 0;
    int rc = time_cfg(cur, cur_ms);
    if (rc) {
        ix->cfg = saved;
        return rc;
    }
    static const uint32_t hs[] = {256, 512, 1024, 2048, 4096}, rgs[] = {4, 8, 16}, mbs[] = {1, 128, 256},
                          wps[] = {1, 2, 4}, scs[] = {1, 2}, wqs[] = {1, 2, 4};
    struct Axis {
        uint32_t wv_search_config::*field;
        const uint32_t *vals;
        int n;
    };
    const Axis axes[] = {{&wv_search_config::visited_slots, hs, 5},
                         {&wv_search_config::row_groups, rgs, 3},
                         {&wv_search_config::merge_buffer, mbs, 3},
                         {&wv_search_config::waves_per_block, wps, 3},
                         {&wv_search_config::schedule, scs, 2},
                         {&wv_search_config::waves_per_query, wqs, 3}};
    for (const Axis &ax : axes) {
        const wv_search_config base = cur;
        for (int i = 0; i < ax.n; ++i) {
            wv_search_config t = base;
            t.*(ax.field) = ax.vals[i];
            double ms = 0;
            const int e = time_cfg(t, ms);
            if (e == WV_EINVAL || e == WV_ENOTSUP) continue;
            if (e) {
                ix->cfg = saved;
                return e;
            }
            if (ms < 0.98 * cur_ms) {  // a move must beat timing noise
                cur = t;
                cur_ms = ms;
            }
        }
    }
    ix->cfg = cur;
    if (best) *best = cur;
    if (best_ms) *best_ms = cur_ms;
    return WV_OK;
}

}  // namespace

extern "C" {

int wv_index_autotune(wv_index *ix, const float *d_queries, uint32_t Q, uint64_t seed, uint32_t L, uint32_t k,
                      const wv_search_ext *ext, wv_search_config *best, double *best_ms) {
    if (!ix) return WV_EINVAL;
    wv_ctx *c = ix->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    if (!d_queries || Q == 0) return fail(c, WV_EINVAL, "autotune needs sample queries");
    uint32_t Lmax = 0;
    int rc = check_search_args(c, ix->N, Q, nullptr, L, k, nullptr, nullptr, k, Lmax);
    if (rc) return rc;
    return autotune_core(ix, d_queries, Q, seed, L, k, nullptr, nullptr, nullptr, k, Lmax, ext, best, best_ms);
}

int wv_index_autotune_budgets(wv_index *ix, const float *d_queries, uint32_t Q, uint64_t seed,
                              const uint32_t *d_L_per_q, const uint32_t *d_k_per_q, uint32_t k_stride,
                              const wv_search_ext *ext, wv_search_config *best, double *best_ms) {
    if (!ix) return WV_EINVAL;
    wv_ctx *c = ix->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    if (!d_queries || Q == 0) return fail(c, WV_EINVAL, "autotune needs sample queries");
    if (!d_L_per_q || !d_k_per_q) return fail(c, WV_EINVAL, "per-query L and k are required");
    WV_HIP(c, hipSetDevice(c->device));
    std::vector<uint32_t> hl(Q), hk(Q);
    WV_HIP(c, hipMemcpy(hl.data(), d_L_per_q, (size_t)Q * 4, hipMemcpyDeviceToHost));
    WV_HIP(c, hipMemcpy(hk.data(), d_k_per_q, (size_t)Q * 4, hipMemcpyDeviceToHost));
    uint32_t Lmax = 0;
    int rc = check_search_args(c, ix->N, Q, nullptr, 0, 0, hl.data(), hk.data(), k_stride, Lmax);
    if (rc) return rc;
    return autotune_core(ix, d_queries, Q, seed, 0, 0, d_L_per_q, d_k_per_q, hl.data(), k_stride, Lmax, ext, best,
                         best_ms);
}

int wv_search_graph_create(wv_index *ix, const float *d_queries, uint32_t Q, uint64_t seed, uint32_t L,
                           uint32_t k, const uint32_t *d_L_per_q, const uint32_t *d_k_per_q, uint32_t k_stride,
                           uint32_t *d_out_ids, float *d_out_dist, uint32_t *d_evals, uint32_t *d_hops,
                           const wv_search_ext *ext, wv_search_graph **out) {
    if (!ix || !out) return WV_EINVAL;
    *out = nullptr;
    wv_ctx *c = ix->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    if (!d_queries || !d_out_ids || Q == 0) return fail(c, WV_EINVAL, "queries/out_ids are null or Q is 0");
    if (ext && (ext->trace || ext->trace_len || ext->settle_hop))
        return fail(c, WV_ENOTSUP, "a search graph records no traces");
    WV_HIP(c, hipSetDevice(c->device));
    std::vector<uint32_t> hl, hk;
    if (d_L_per_q) {
        hl.resize(Q);
        WV_HIP(c, hipMemcpy(hl.data(), d_L_per_q, (size_t)Q * 4, hipMemcpyDeviceToHost));
    }
    if (d_k_per_q) {
        hk.resize(Q);
        WV_HIP(c, hipMemcpy(hk.data(), d_k_per_q, (size_t)Q * 4, hipMemcpyDeviceToHost));
    }
    uint32_t Lmax = 0;
    int rc = check_search_args(c, ix->N, Q, nullptr, L, k, d_L_per_q ? hl.data() : nullptr,
                               d_k_per_q ? hk.data() : nullptr, k_stride, Lmax);
    if (rc) return rc;
    std::unique_ptr<wv_search_graph> sg(new wv_search_graph);
    sg->ix = ix;
    SearchExtDev x;
    if (ext && ext->model) {
        WV_HIP(c, hipMalloc(&sg->model, sizeof(wv_term_model)));
        WV_HIP(c, hipMemcpy(sg->model, ext->model, sizeof(wv_term_model), hipMemcpyHostToDevice));
        x.model = static_cast<const wv_term_model *>(sg->model);
    }
    wv::SearchArgs a{};
    SearchPlan pl;
    int cpl = 0;
    rc = prepare_search(ix, d_queries, Q, nullptr, seed, L, k, d_L_per_q, d_k_per_q, Lmax, k_stride, d_out_ids,
                        d_out_dist, d_evals, d_hops, x, a, pl, cpl);
    if (!rc) rc = attach_queue(ix, pl, d_L_per_q ? hl.data() : nullptr, Q, sg->order, a);
    if (rc) {
        if (sg->model) (void)hipFree(sg->model);
        return rc;
    }
    // one eager launch sets the kernel's LDS attribute outside the capture and
    // fills the outputs once (the persistent schedule's counter is zeroed by
    // a kernel inside the launch, so every replay starts the queue afresh)
    hipError_t e = launch_plan(a, pl, cpl, c);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        hipError_t le = launch_plan(a, pl, cpl, c);
        e = hipStreamEndCapture(c->stream, &sg->graph);
        if (e == hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&sg->exec, sg->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        if (sg->exec) (void)hipGraphExecDestroy(sg->exec);
        if (sg->graph) (void)hipGraphDestroy(sg->graph);
        if (sg->model) (void)hipFree(sg->model);
        return fail(c, WV_EHIP, "search graph capture: %s", hipGetErrorString(e));
    }
    *out = sg.release();
    return WV_OK;
}

int wv_search_graph_launch(wv_search_graph *sg, int sync) {
    if (!sg) return WV_EINVAL;
    wv_ctx *c = sg->ix->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    WV_HIP(c, hipSetDevice(c->device));
    WV_HIP(c, hipGraphLaunch(sg->exec, c->stream));
    if (sync) WV_HIP(c, hipStreamSynchronize(c->stream));
    return WV_OK;
}

int wv_search_graph_destroy(wv_search_graph *sg) {
    if (!sg) return WV_OK;
    wv_ctx *c = sg->ix->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (sg->exec) (void)hipGraphExecDestroy(sg->exec);
    if (sg->graph) (void)hipGraphDestroy(sg->graph);
    if (sg->model) (void)hipFree(sg->model);
    delete sg;
    return WV_OK;
}

int wv_mlp_grad(wv_ctx *c, const wv_term_model *m, const float *X, const float *y, uint32_t n,
                float pos_weight, float *grad, double *loss) {
    if (!c) return WV_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (!m || !X || !y || !grad || n == 0) return fail(c, WV_EINVAL, "null argument or no samples");
    WV_HIP(c, hipSetDevice(c->device));
    const uint32_t nb = (n + 63) / 64;
    DevBuf dp, dx, dy, dpart, dg, dl;
    WV_HIP(c, dp.alloc(sizeof(wv_term_model)));
    WV_HIP(c, dx.alloc((size_t)n * 8 * 4));
    WV_HIP(c, dy.alloc((size_t)n * 4));
    WV_HIP(c, dpart.alloc((size_t)nb * wv::MLP_PARTIAL_STRIDE * 4));
    WV_HIP(c, dg.alloc(wv::MLP_PARAMS * 4));
    WV_HIP(c, dl.alloc(16));
    WV_HIP(c, hipMemcpyAsync(dp.p, m, sizeof(wv_term_model), hipMemcpyHostToDevice, c->stream));
    WV_HIP(c, hipMemcpyAsync(dx.p, X, (size_t)n * 8 * 4, hipMemcpyHostToDevice, c->stream));
    WV_HIP(c, hipMemcpyAsync(dy.p, y, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    const float *params = dp.as<float>();
    WV_HIP(c, wv::mlp_grad_launch(params, model_field(dp.p, offsetof(wv_term_model, mean)),
                                  model_field(dp.p, offsetof(wv_term_model, inv_std)), dx.as<float>(),
                                  dy.as<float>(), nullptr, n,
                                  pos_weight, dpart.as<float>(), c->stream));
    WV_HIP(c, wv::mlp_update_launch(dpart.as<float>(), nb, 1.0f / n, nullptr, nullptr, nullptr, dg.as<float>(),
                                    dl.as<float>(), 0, 0.f, 0.f, 0.f, 0.f, 1.f, 1.f, c->stream));
    WV_HIP(c, hipStreamSynchronize(c->stream));
    WV_HIP(c, hipMemcpy(grad, dg.p, wv::MLP_PARAMS * 4, hipMemcpyDeviceToHost));
    if (loss) {
        float l = 0.f;
        WV_HIP(c, hipMemcpy(&l, dl.p, 4, hipMemcpyDeviceToHost));
        *loss = l;
    }
    return WV_OK;
}

int wv_mlp_train(wv_ctx *c, const float *X, const float *y, uint32_t n, const wv_train_params *tp,
                 wv_term_model *model, wv_train_stats *stats) {
    if (!c) return WV_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (!X || !y || !tp || !model || n == 0) return fail(c, WV_EINVAL, "null argument or no samples");
    if (tp->batch == 0 || tp->batch % 64) return fail(c, WV_EINVAL, "batch must be a positive multiple of 64");
    WV_HIP(c, hipSetDevice(c->device));
    // standardisation from the data unless the model carries one
    bool have_std = false;
    for (int i = 0; i < 8; ++i) have_std |= model->inv_std[i] != 0.f;
    if (!have_std) {
        for (int i = 0; i < 8; ++i) {
            double s = 0, s2 = 0;
            for (uint32_t r = 0; r < n; ++r) {
                const double v = X[(size_t)r * 8 + i];
                s += v;
                s2 += v * v;
            }
            const double mu = s / n, var = std::max(s2 / n - mu * mu, 0.0);
            model->mean[i] = (float)mu;
            model->inv_std[i] = var > 1e-20 ? (float)(1.0 / std::sqrt(var)) : 1.0f;
        }
    }
    const uint32_t B = tp->batch;
    const uint32_t steps_per_epoch = (n + B - 1) / B;
    const uint32_t nbmax = B / 64;
    DevBuf dp, dx, dy, didx, dpart, dmo, dvo, dloss;
    WV_HIP(c, dp.alloc(sizeof(wv_term_model)));
    WV_HIP(c, dx.alloc((size_t)n * 8 * 4));
    WV_HIP(c, dy.alloc((size_t)n * 4));
    WV_HIP(c, didx.alloc((size_t)n * 4));
    WV_HIP(c, dpart.alloc((size_t)nbmax * wv::MLP_PARTIAL_STRIDE * 4));
    WV_HIP(c, dmo.alloc(wv::MLP_PARAMS * 4));
    WV_HIP(c, dvo.alloc(wv::MLP_PARAMS * 4));
    const size_t total_steps = (size_t)tp->epochs * steps_per_epoch;
    WV_HIP(c, dloss.alloc(std::max<size_t>(total_steps, 1) * 4));
    WV_HIP(c, hipMemcpyAsync(dp.p, model, sizeof(wv_term_model), hipMemcpyHostToDevice, c->stream));
    WV_HIP(c, hipMemcpyAsync(dx.p, X, (size_t)n * 8 * 4, hipMemcpyHostToDevice, c->stream));
    WV_HIP(c, hipMemcpyAsync(dy.p, y, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    WV_HIP(c, hipMemsetAsync(dmo.p, 0, wv::MLP_PARAMS * 4, c->stream));
    WV_HIP(c, hipMemsetAsync(dvo.p, 0, wv::MLP_PARAMS * 4, c->stream));
    float *params = dp.as<float>();
    const float *dmean = model_field(dp.p, offsetof(wv_term_model, mean));
    const float *dinv = model_field(dp.p, offsetof(wv_term_model, inv_std));
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    std::mt19937_64 rng(tp->seed);
    WV_HIP(c, hipEventRecord(c->ev0, c->stream));
    size_t step = 0;
    std::vector<uint32_t> batch_n(total_steps);
    for (uint32_t e = 0; e < tp->epochs; ++e) {
        std::shuffle(perm.begin(), perm.end(), rng);
        WV_HIP(c, hipStreamSynchronize(c->stream));  // perm buffer reuse
        WV_HIP(c, hipMemcpyAsync(didx.p, perm.data(), (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
        for (uint32_t b0 = 0; b0 < n; b0 += B, ++step) {
            const uint32_t nbt = std::min(B, n - b0);
            batch_n[step] = nbt;
            const float bc1 = 1.f - std::pow(tp->beta1, (float)(step + 1));
            const float bc2 = 1.f - std::pow(tp->beta2, (float)(step + 1));
            WV_HIP(c, wv::mlp_grad_launch(params, dmean, dinv, dx.as<float>(), dy.as<float>(),
                                          didx.as<uint32_t>() + b0, nbt, tp->pos_weight, dpart.as<float>(),
                                          c->stream));
            WV_HIP(c, wv::mlp_update_launch(dpart.as<float>(), (nbt + 63) / 64, 1.0f / nbt, params, dmo.as<float>(),
                                            dvo.as<float>(), nullptr, dloss.as<float>() + step, 1, tp->lr,
                                            tp->beta1, tp->beta2, tp->eps, bc1, bc2, c->stream));
        }
    }
    WV_HIP(c, hipEventRecord(c->ev1, c->stream));
    WV_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<float> losses(total_steps);
    if (total_steps) WV_HIP(c, hipMemcpy(losses.data(), dloss.p, total_steps * 4, hipMemcpyDeviceToHost));
    WV_HIP(c, hipMemcpy(model, dp.p, wv::MLP_PARAMS * 4, hipMemcpyDeviceToHost));  // w1, b1, w2, b2
    if (stats) {
        float ms = 0.f;
        WV_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        stats->kernel_ms = ms;
        stats->steps = (uint32_t)total_steps;
        auto epoch_loss = [&](uint32_t e) {
            double s = 0;
            for (uint32_t t = 0; t < steps_per_epoch; ++t) {
                const size_t i = (size_t)e * steps_per_epoch + t;
                s += (double)losses[i] * batch_n[i];
            }
            return s / n;
        };
        stats->first_loss = tp->epochs ? epoch_loss(0) : 0.0;
        stats->last_loss = tp->epochs ? epoch_loss(tp->epochs - 1) : 0.0;
    }
    return WV_OK;
}

int wv_mlp_predict(wv_ctx *c, const wv_term_model *m, const float *X, uint32_t n, float *logits) {
    if (!c) return WV_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (!m || !X || !logits) return fail(c, WV_EINVAL, "null argument");
    if (n == 0) return WV_OK;
    WV_HIP(c, hipSetDevice(c->device));
    DevBuf dm, dx, dz;
    WV_HIP(c, dm.alloc(sizeof(wv_term_model)));
    WV_HIP(c, dx.alloc((size_t)n * 8 * 4));
    WV_HIP(c, dz.alloc((size_t)n * 4));
    WV_HIP(c, hipMemcpyAsync(dm.p, m, sizeof(wv_term_model), hipMemcpyHostToDevice, c->stream));
    WV_HIP(c, hipMemcpyAsync(dx.p, X, (size_t)n * 8 * 4, hipMemcpyHostToDevice, c->stream));
    WV_HIP(c, wv::mlp_predict_launch(dm.as<wv_term_model>(), dx.as<float>(), n, dz.as<float>(), c->stream));
    WV_HIP(c, hipStreamSynchronize(c->stream));
    WV_HIP(c, hipMemcpy(logits, dz.p, (size_t)n * 4, hipMemcpyDeviceToHost));
    return WV_OK;
}

namespace {
int check_nnd_params(wv_ctx *c, const wv_index *ix, const wv_nnd_params *p) {
    if (!p) return fail(c, WV_EINVAL, "params are null");
    if (p->K == 0 || p->L < p->K)
        return fail(c, WV_EINVAL, "need 1 <= K <= L (K=%u L=%u)", p->K, p->L);
    if (p->S == 0 || p->S >= ix->N || p->S > p->L)
        return fail(c, WV_EINVAL, "need 1 <= S < N and S <= L (S=%u N=%u L=%u)", p->S, ix->N, p->L);
    if (p->S > 64) return fail(c, WV_ENOTSUP, "S above 64");
    if (p->L > 1024 || p->R > 1024) return fail(c, WV_ENOTSUP, "L or R above 1024");
    if (p->dtype > WV_DTYPE_FP8) return fail(c, WV_EINVAL, "unknown dtype %u", p->dtype);
    if (p->cand_slots & (p->cand_slots - 1)) return fail(c, WV_EINVAL, "cand_slots must be a power of two");
    return WV_OK;
}

int run_nnd(wv_index *ix, const wv_nnd_params *p, const uint32_t *init_ids, const uint32_t *init_cnt,
            wv_build_stats *stats, bool efanna, uint32_t init_stride, const wv::NndMonitor *monitor = nullptr) {
    wv_ctx *c = ix->ctx;
    WV_HIP(c, hipSetDevice(c->device));
    int rc = wv::index_alloc_graph(ix, p->K);
    if (rc) return rc;
    rc = wv::nndescent_build(ix, p, init_ids, init_cnt, stats, ix->graph, ix->gdist, ix->deg, efanna, init_stride,
                             monitor);
    if (rc) {
        (void)hipFree(ix->graph);
        (void)hipFree(ix->gdist);
        (void)hipFree(ix->deg);
        ix->graph = nullptr;
        ix->gdist = nullptr;
        ix->deg = nullptr;
        ix->deg_max = 0;
    }
    return rc;
}
}  // namespace

int wv_index_build_nndescent(wv_index *ix, const wv_nnd_params *p, const uint32_t *init_ids,
                             const uint32_t *init_cnt, wv_build_stats *stats) {
    if (!ix) return WV_EINVAL;
    wv_ctx *c = ix->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    int rc = check_nnd_params(c, ix, p);
    if (rc) return rc;
    return run_nnd(ix, p, init_ids, init_cnt, stats, false, 0);
}

int wv_index_build_nndescent_monitor(wv_index *ix, const wv_nnd_params *p, const uint32_t *init_ids,
                                     const uint32_t *init_cnt, const uint32_t *ctrl, uint32_t n_ctrl,
                                     const uint32_t *truth, uint32_t truth_k, double *per_round,
                                     wv_build_stats *stats) {
    if (!ix) return WV_EINVAL;
    wv_ctx *c = ix->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    int rc = check_nnd_params(c, ix, p);
    if (rc) return rc;
    if (!ctrl || !truth || !per_round) return fail(c, WV_EINVAL, "ctrl/truth/per_round are null");
    if (n_ctrl == 0 || n_ctrl > (1u << 20) || truth_k == 0 || truth_k > 1024)
        return fail(c, WV_EINVAL, "need 1 <= n_ctrl <= 2^20 and 1 <= truth_k <= 1024 (n_ctrl=%u truth_k=%u)",
                    n_ctrl, truth_k);
    for (uint32_t i = 0; i < n_ctrl; ++i)
        if (ctrl[i] >= ix->N) return fail(c, WV_EINVAL, "control node %u is %u, not below N=%u", i, ctrl[i], ix->N);
    WV_HIP(c, hipSetDevice(c->device));
    const uint32_t rounds = std::max<uint32_t>(p->iters, 1);
    DevBuf dc, dt, dh;
    WV_HIP(c, dc.alloc((size_t)n_ctrl * 4));
    WV_HIP(c, dt.alloc((size_t)n_ctrl * truth_k * 4));
    WV_HIP(c, dh.alloc((size_t)rounds * n_ctrl * 4));
    WV_HIP(c, hipMemcpyAsync(dc.p, ctrl, (size_t)n_ctrl * 4, hipMemcpyHostToDevice, c->stream));
    WV_HIP(c, hipMemcpyAsync(dt.p, truth, (size_t)n_ctrl * truth_k * 4, hipMemcpyHostToDevice, c->stream));
    WV_HIP(c, hipMemsetAsync(dh.p, 0, (size_t)rounds * n_ctrl * 4, c->stream));
    wv::NndMonitor m{dc.as<uint32_t>(), dt.as<uint32_t>(), n_ctrl, truth_k, dh.as<uint32_t>()};
    rc = run_nnd(ix, p, init_ids, init_cnt, stats, false, 0, &m);
    if (rc) return rc;
    std::vector<uint32_t> hits((size_t)rounds * n_ctrl);
    WV_HIP(c, hipMemcpy(hits.data(), dh.p, hits.size() * 4, hipMemcpyDeviceToHost));
    for (uint32_t r = 0; r < p->iters; ++r) {
        double sum = 0.0;
        for (uint32_t i = 0; i < n_ctrl; ++i) sum += (double)hits[(size_t)r * n_ctrl + i] / truth_k;
        per_round[r] = sum / n_ctrl;
    }
    return WV_OK;
}

int wv_index_build_efanna(wv_index *ix, const wv_nnd_params *p, const uint32_t *init_ids,
                          const uint32_t *init_cnt, uint32_t init_stride, wv_build_stats *stats) {
    if (!ix) return WV_EINVAL;
    wv_ctx *c = ix->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    int rc = check_nnd_params(c, ix, p);
    if (rc) return rc;
    if (2ull * p->S >= ix->N) return fail(c, WV_EINVAL, "EFANNA samples 2S ids: need 2S < N (S=%u N=%u)", p->S, ix->N);
    if (init_ids && (init_stride == 0 || init_stride > p->L))
        return fail(c, WV_EINVAL, "need 1 <= init_stride <= L (init_stride=%u L=%u)", init_stride, p->L);
    return run_nnd(ix, p, init_ids, init_cnt, stats, true, init_stride);
}

int wv_index_build_nsg(wv_index *ix, const wv_nsg_params *p, wv_nsg_stats *stats) {
    if (!ix) return WV_EINVAL;
    wv_ctx *c = ix->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    if (!p) return fail(c, WV_EINVAL, "params are null");
    if (!ix->graph) return fail(c, WV_EINVAL, "index has no graph (build or set the kNN graph first)");
    if (p->L == 0 || p->L >= ix->N) return fail(c, WV_EINVAL, "need 1 <= L < N (L=%u N=%u)", p->L, ix->N);
    if (p->L > 1024 || p->C == 0 || p->C > 1024)
        return fail(c, WV_ENOTSUP, "need L <= 1024 and 1 <= C <= 1024 (L=%u C=%u)", p->L, p->C);
    if (p->R == 0 || p->R > 256) return fail(c, WV_ENOTSUP, "need 1 <= R <= 256 (R=%u)", p->R);
    if (p->ep >= ix->N) return fail(c, WV_EINVAL, "ep %u out of range", p->ep);
    if (ix->dp > 1024) return fail(c, WV_ENOTSUP, "dimension above 1024");
    if ((uint64_t)ix->N * p->R >= (1ull << 31))
        return fail(c, WV_ENOTSUP, "N x R above 2^31 edges (the edge offsets are 32-bit)");
    WV_HIP(c, hipSetDevice(c->device));
    // the kNN graph stays alive until the result replaces it
    return wv::nsg_build(ix, p, stats, ix->graph, ix->deg, ix->deg_max, wv::RefineRule{true, false, true, 1.0f});
}

int wv_index_build_ssg(wv_index *ix, const wv_ssg_params *p, uint32_t *eps, wv_nsg_stats *stats) {
    if (!ix) return WV_EINVAL;
    wv_ctx *c = ix->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    if (!p) return fail(c, WV_EINVAL, "params are null");
    if (!ix->graph) return fail(c, WV_EINVAL, "index has no graph (build or set the kNN graph first)");
    if (p->L == 0 || p->L > 2048) return fail(c, WV_ENOTSUP, "need 1 <= L <= 2048 (L=%u)", p->L);
    if (p->R == 0 || p->R > 256) return fail(c, WV_ENOTSUP, "need 1 <= R <= 256 (R=%u)", p->R);
    if (p->n_try == 0 || p->n_try >= ix->N) return fail(c, WV_EINVAL, "need 1 <= n_try < N (n_try=%u)", p->n_try);
    if (!(p->angle > 0.0f) || !(p->angle < 180.0f)) return fail(c, WV_EINVAL, "angle must be in (0, 180) degrees");
    if (ix->dp > 1024) return fail(c, WV_ENOTSUP, "dimension above 1024");
    if ((uint64_t)ix->N * p->R >= (1ull << 31))
        return fail(c, WV_ENOTSUP, "N x R above 2^31 edges (the edge offsets are 32-bit)");
    WV_HIP(c, hipSetDevice(c->device));
    return wv::ssg_build(ix, p, eps, stats, ix->graph, ix->deg, ix->deg_max);
}

int wv_index_build_dpg(wv_index *ix, uint32_t range, wv_nsg_stats *stats) {
    if (!ix) return WV_EINVAL;
    wv_ctx *c = ix->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    if (!ix->graph) return fail(c, WV_EINVAL, "index has no graph (build or set the kNN graph first)");
    if (range == 0 || range > 256) return fail(c, WV_ENOTSUP, "need 1 <= range <= 256 (range=%u)", range);
    if (ix->dp > 1024) return fail(c, WV_ENOTSUP, "dimension above 1024");
    if ((uint64_t)ix->N * range >= (1ull << 31))
        return fail(c, WV_ENOTSUP, "N x range above 2^31 edges (the edge offsets are 32-bit)");
    WV_HIP(c, hipSetDevice(c->device));
    return wv::dpg_build(ix, range, stats, ix->graph, ix->deg, ix->deg_max);
}

int wv_index_build_vamana(wv_index *ix, const wv_vamana_params *p, wv_nsg_stats *stats) {
    if (!ix) return WV_EINVAL;
    wv_ctx *c = ix->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    if (!p) return fail(c, WV_EINVAL, "params are null");
    if (!ix->graph) return fail(c, WV_EINVAL, "index has no graph (build or set the initial graph first)");
    if (p->L == 0 || p->L >= ix->N) return fail(c, WV_EINVAL, "need 1 <= L < N (L=%u N=%u)", p->L, ix->N);
    if (p->L > 1024 || p->C == 0 || p->C > 1024)
        return fail(c, WV_ENOTSUP, "need L <= 1024 and 1 <= C <= 1024 (L=%u C=%u)", p->L, p->C);
    if (p->R == 0 || p->R > 256) return fail(c, WV_ENOTSUP, "need 1 <= R <= 256 (R=%u)", p->R);
    if (p->ep >= ix->N) return fail(c, WV_EINVAL, "ep %u out of range", p->ep);
    if (!(p->alpha >= 1.0f) || !(p->alpha < 1e30f)) return fail(c, WV_EINVAL, "alpha must be finite and >= 1");
    if (ix->dp > 1024) return fail(c, WV_ENOTSUP, "dimension above 1024");
    if ((uint64_t)ix->N * p->R >= (1ull << 31))
        return fail(c, WV_ENOTSUP, "N x R above 2^31 edges (the edge offsets are 32-bit)");
    WV_HIP(c, hipSetDevice(c->device));
    wv_nsg_params q{p->L, p->R, p->C, p->ep, p->seed};
    return wv::nsg_build(ix, &q, stats, ix->graph, ix->deg, ix->deg_max, wv::RefineRule{false, true, false, p->alpha});
}

int wv_nndescent_build(wv_ctx *c, const float *base, uint32_t N, uint32_t d,
                       const wv_nnd_params *p, const uint32_t *init_ids, uint32_t *out_ids,
                       float *out_dist, uint32_t *out_deg, wv_build_stats *stats) {
    if (!c) return WV_EINVAL;
    if (!out_ids || !out_deg) return fail(c, WV_EINVAL, "out_ids/out_deg are null");
    wv_index *ix = nullptr;
    int rc = wv_index_create(c, base, N, d, &ix);
    if (rc) return rc;
    rc = wv_index_build_nndescent(ix, p, init_ids, nullptr, stats);
    if (!rc) rc = wv_index_get_graph(ix, out_ids, out_dist, out_deg, nullptr);
    wv_index_destroy(ix);
    return rc;
}

int wv_search_greedy(wv_ctx *c, const wv_graph *gr, const float *base, uint32_t d,
                     const float *queries, uint32_t Q, const uint32_t *entries, uint64_t seed,
                     uint32_t L, uint32_t k, const uint32_t *L_per_q, const uint32_t *k_per_q,
                     uint32_t k_stride, uint32_t *out_ids, float *out_dist, uint32_t *evals,
                     uint32_t *hops, wv_search_stats *stats) {
    if (!c) return WV_EINVAL;
    if (!gr) return fail(c, WV_EINVAL, "graph is null");
    wv_index *ix = nullptr;
    int rc = wv_index_create(c, base, gr->N, d, &ix);
    if (rc) return rc;
    rc = wv_index_set_graph(ix, gr->ids, gr->deg, gr->deg_max);
    if (!rc)
        rc = wv_index_search(ix, queries, Q, entries, seed, L, k, L_per_q, k_per_q, k_stride,
                             out_ids, out_dist, evals, hops, stats);
    wv_index_destroy(ix);
    return rc;
}

int wv_gather_rate(wv_ctx *c, uint32_t N, uint32_t d, uint32_t waves_per_cu, uint32_t batches, double *gbps,
                   double *kernel_ms) {
    if (!c) return WV_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (!gbps) return fail(c, WV_EINVAL, "null argument");
    if (N == 0 || batches == 0) return fail(c, WV_EINVAL, "N and batches must be positive");
    if (d == 0 || d > 256) return fail(c, WV_ENOTSUP, "gather rate supports 1 <= d <= 256");
    if (waves_per_cu == 0 || waves_per_cu > 16) return fail(c, WV_EINVAL, "waves_per_cu must be 1..16");
    if (c->cu_count <= 0) return fail(c, WV_EHIP, "compute unit count unknown");
    WV_HIP(c, hipSetDevice(c->device));
    const uint32_t dp = (d + 3) / 4 * 4;
    const uint32_t blocks = (uint32_t)c->cu_count * waves_per_cu;
    // LDS a workgroup, in whole KiB, so that exactly waves_per_cu fit the
    // 160 KiB of a CU (at least the 768 bytes the kernel uses)
    const uint32_t lds = std::max<uint32_t>(768u, (160u * 1024u / waves_per_cu) & ~1023u);
    DevBuf db, dout;
    WV_HIP(c, db.alloc((size_t)N * dp * 4));
    WV_HIP(c, dout.alloc((size_t)blocks * 64 * 8));
    WV_HIP(c, hipMemsetAsync(db.p, 0, (size_t)N * dp * 4, c->stream));
    // one untimed launch, then the timed one
    WV_HIP(c, wv::launch_gather_rate(db.as<float>(), N, dp, blocks, std::min<uint32_t>(batches, 4u), lds,
                                     dout.as<uint64_t>(), c->stream));
    WV_HIP(c, hipEventRecord(c->ev0, c->stream));
    WV_HIP(c, wv::launch_gather_rate(db.as<float>(), N, dp, blocks, batches, lds, dout.as<uint64_t>(), c->stream));
    WV_HIP(c, hipEventRecord(c->ev1, c->stream));
    WV_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    WV_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (kernel_ms) *kernel_ms = ms;
    *gbps = (double)blocks * 64.0 * batches * 4.0 * d / (ms * 1e-3) / 1e9;
    return WV_OK;
}

int wv_line_rate(wv_ctx *c, uint64_t bytes, uint32_t iters, double *gaccess, double *kernel_ms) {
    if (!c) return WV_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (!gaccess) return fail(c, WV_EINVAL, "null argument");
    if (bytes < ((uint64_t)1 << 20) || bytes % 16) return fail(c, WV_EINVAL, "bytes must be a multiple of 16, >= 1 MiB");
    if (iters == 0) return fail(c, WV_EINVAL, "iters must be positive");
    if (c->cu_count <= 0) return fail(c, WV_EHIP, "compute unit count unknown");
    WV_HIP(c, hipSetDevice(c->device));
    // 32 workgroups of 256 a CU: the 8 wavefronts a SIMD the hardware holds
    const uint32_t blocks = (uint32_t)c->cu_count * 32;
    DevBuf db, dout;
    WV_HIP(c, db.alloc(bytes));
    WV_HIP(c, dout.alloc((size_t)blocks * 256 * 4));
    WV_HIP(c, hipMemsetAsync(db.p, 0, bytes, c->stream));
    // one short untimed launch, then the timed one
    WV_HIP(c, wv::launch_line_rate(db.p, bytes, blocks, std::min<uint32_t>(iters, 4u), dout.as<uint32_t>(), c->stream));
    WV_HIP(c, hipEventRecord(c->ev0, c->stream));
    WV_HIP(c, wv::launch_line_rate(db.p, bytes, blocks, iters, dout.as<uint32_t>(), c->stream));
    WV_HIP(c, hipEventRecord(c->ev1, c->stream));
    WV_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    WV_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (kernel_ms) *kernel_ms = ms;
    *gaccess = (double)blocks * 256.0 * iters * 4.0 / (ms * 1e-3) / 1e9;
    return WV_OK;
}

int wv_exact_knn(wv_ctx *c, const float *base, uint32_t N, const float *queries, uint32_t Q,
                 uint32_t d, uint32_t k, int64_t self_offset, uint32_t *out_ids, double *out_dist,
                 double *kernel_ms) {
    if (!c) return WV_EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    if (!base || !queries || !out_ids || !out_dist) return fail(c, WV_EINVAL, "null argument");
    if (k == 0 || k > 128) return fail(c, WV_ENOTSUP, "exact kNN supports 1 <= k <= 128 (k=%u)", k);
    if (k > N - (self_offset >= 0 ? 1u : 0u)) return fail(c, WV_EINVAL, "k above the number of candidates");
    if (d == 0 || d > 1024) return fail(c, WV_ENOTSUP, "dimension must be 1..1024");
    WV_HIP(c, hipSetDevice(c->device));
    if (Q == 0) return WV_OK;
    DevBuf db, dq, di, dd;
    WV_HIP(c, db.alloc((size_t)N * d * 4));
    WV_HIP(c, dq.alloc((size_t)Q * d * 4));
    WV_HIP(c, di.alloc((size_t)Q * k * 4));
    WV_HIP(c, dd.alloc((size_t)Q * k * 8));
    WV_HIP(c, hipMemcpyAsync(db.p, base, (size_t)N * d * 4, hipMemcpyHostToDevice, c->stream));
    WV_HIP(c, hipMemcpyAsync(dq.p, queries, (size_t)Q * d * 4, hipMemcpyHostToDevice, c->stream));
    wv::ExactArgs a;
    a.base = db.as<float>();
    a.queries = dq.as<float>();
    a.N = N;
    a.Q = Q;
    a.d = d;
    a.ld_base = d;
    a.ld_query = d;
    a.k = k;
    a.self_offset = self_offset;
    a.out_ids = di.as<uint32_t>();
    a.out_dist = dd.as<double>();
    WV_HIP(c, hipEventRecord(c->ev0, c->stream));
    WV_HIP(c, wv::launch_exact_knn(a, c->stream));
    WV_HIP(c, hipEventRecord(c->ev1, c->stream));
    WV_HIP(c, hipStreamSynchronize(c->stream));
    if (kernel_ms) {
        float ms = 0.f;
        WV_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        *kernel_ms = ms;
    }
    WV_HIP(c, hipMemcpy(out_ids, di.p, (size_t)Q * k * 4, hipMemcpyDeviceToHost));
    WV_HIP(c, hipMemcpy(out_dist, dd.p, (size_t)Q * k * 8, hipMemcpyDeviceToHost));
    return WV_OK;
}

}  // extern "C"
