"""ctypes binding of libweavess_hip.so (the C ABI in include/weavess_hip.h).

The library is built in-tree by ``weavess_amd.build``; importing this module
does not build anything and raises if the library is missing, so a GPU run can
never fall back to another implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libweavess_hip.so")

WV_OK = 0
WV_EINVAL = -22
WV_ENOMEM = -12
WV_EHIP = -5
WV_ENOTSUP = -95

DTYPES = {"f32": 0, "bf16": 1, "fp8": 2}

# every symbol declared in include/weavess_hip.h
EXPORTS = [
    "wv_version", "wv_ctx_create", "wv_ctx_destroy", "wv_last_error",
    "wv_index_create", "wv_index_destroy", "wv_index_set_graph", "wv_index_get_graph",
    "wv_index_set_search_config", "wv_index_build_nndescent", "wv_index_search",
    "wv_index_search_dev", "wv_nndescent_build", "wv_search_greedy", "wv_exact_knn",
    "wv_index_search_ext", "wv_index_search_ext_dev", "wv_index_autotune", "wv_search_graph_create",
    "wv_search_graph_launch", "wv_search_graph_destroy", "wv_mlp_grad", "wv_mlp_train",
    "wv_mlp_predict", "wv_index_build_efanna", "wv_index_build_nsg", "wv_index_build_vamana",
    "wv_index_build_dpg", "wv_index_build_ssg", "wv_index_search_dev_noalloc",
    "wv_index_build_nndescent_monitor", "wv_gather_rate", "wv_line_rate", "wv_index_autotune_budgets",
]
WV_NOALLOC_MAX_Q = 1 << 18


class WvError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"weavess_hip error {code}: {msg}")
        self.code = code


class NndParams(C.Structure):
    _fields_ = [("K", C.c_uint32), ("L", C.c_uint32), ("S", C.c_uint32), ("R", C.c_uint32),
                ("iters", C.c_uint32), ("dtype", C.c_uint32), ("rerank_mult", C.c_uint32),
                ("cand_slots", C.c_uint32), ("seed", C.c_uint64), ("order", C.c_uint32),
                ("reserved", C.c_uint32)]


class BuildStats(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("init_ms", C.c_double), ("sample_ms", C.c_double),
                ("join_ms", C.c_double), ("merge_ms", C.c_double), ("finalize_ms", C.c_double),
                ("pairs", C.c_uint64), ("proposals", C.c_uint64), ("rows_gathered", C.c_uint64),
                ("iters_run", C.c_uint32), ("cand_slots", C.c_uint32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class NsgParams(C.Structure):
    _fields_ = [("L", C.c_uint32), ("R", C.c_uint32), ("C", C.c_uint32), ("ep", C.c_uint32),
                ("seed", C.c_uint64)]


class VamanaParams(C.Structure):
    _fields_ = [("L", C.c_uint32), ("R", C.c_uint32), ("C", C.c_uint32), ("ep", C.c_uint32),
                ("alpha", C.c_float), ("reserved", C.c_uint32), ("seed", C.c_uint64)]


class SsgParams(C.Structure):
    _fields_ = [("L", C.c_uint32), ("R", C.c_uint32), ("angle", C.c_float), ("n_try", C.c_uint32),
                ("seed", C.c_uint64)]


class NsgStats(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("search_ms", C.c_double), ("prune_ms", C.c_double),
                ("insert_ms", C.c_double), ("conn_ms", C.c_double), ("evals", C.c_uint64),
                ("conn_edges", C.c_uint32), ("width", C.c_uint32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SearchStats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("dist_evals", C.c_uint64), ("hops", C.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Graph(C.Structure):
    _fields_ = [("ids", C.c_void_p), ("deg", C.c_void_p), ("N", C.c_uint32),
                ("deg_max", C.c_uint32)]


class SearchConfig(C.Structure):
    _fields_ = [("waves_per_block", C.c_uint32), ("row_groups", C.c_uint32),
                ("visited_slots", C.c_uint32), ("merge_buffer", C.c_uint32), ("schedule", C.c_uint32),
                ("waves_per_query", C.c_uint32)]


class TermModel(C.Structure):
    """wv_term_model: 8 -> 32 (ReLU) -> 1 network, standardisation, stop rule."""
    _fields_ = [("w1", (C.c_float * 8) * 32), ("b1", C.c_float * 32), ("w2", C.c_float * 32),
                ("b2", C.c_float), ("mean", C.c_float * 8), ("inv_std", C.c_float * 8),
                ("logit_threshold", C.c_float), ("min_hops", C.c_uint32)]

    def params(self) -> np.ndarray:
        """The 321 trainable parameters (w1, b1, w2, b2) as float32."""
        return np.frombuffer(bytes(self), np.float32, 321).copy()

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == 321
        C.memmove(C.addressof(self), p.ctypes.data, 321 * 4)

    @classmethod
    def init(cls, seed=0):
        """He-initialised weights, zero biases, no standardisation yet."""
        g = np.random.default_rng(seed)
        m = cls()
        p = np.zeros(321, np.float32)
        p[:256] = g.normal(0, np.sqrt(2.0 / 8), 256)
        p[288:320] = g.normal(0, np.sqrt(1.0 / 32), 32)
        m.set_params(p)
        m.logit_threshold = 1e30
        return m


class SearchExt(C.Structure):
    _fields_ = [("model", C.c_void_p), ("trace", C.c_void_p), ("trace_len", C.c_void_p),
                ("settle_hop", C.c_void_p), ("trace_cap", C.c_uint32)]


class TrainParams(C.Structure):
    _fields_ = [("epochs", C.c_uint32), ("batch", C.c_uint32), ("lr", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("pos_weight", C.c_float), ("seed", C.c_uint64)]


class TrainStats(C.Structure):
    _fields_ = [("first_loss", C.c_double), ("last_loss", C.c_double), ("kernel_ms", C.c_double),
                ("steps", C.c_uint32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None


def load() -> C.CDLL:
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m weavess_amd.build` first")
    L = C.CDLL(LIB_PATH)
    vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    sig = {
        "wv_version": ([], i32),
        "wv_ctx_create": ([i32, C.POINTER(vp)], i32),
        "wv_ctx_destroy": ([vp], i32),
        "wv_last_error": ([vp], C.c_char_p),
        "wv_index_create": ([vp, vp, u32, u32, C.POINTER(vp)], i32),
        "wv_index_destroy": ([vp], i32),
        "wv_index_set_graph": ([vp, vp, vp, u32], i32),
        "wv_index_get_graph": ([vp, vp, vp, vp, vp], i32),
        "wv_index_set_search_config": ([vp, vp], i32),
        "wv_index_build_nndescent": ([vp, vp, vp, vp, vp], i32),
        "wv_index_build_efanna": ([vp, vp, vp, vp, u32, vp], i32),
        "wv_index_build_nndescent_monitor": ([vp, vp, vp, vp, vp, u32, vp, u32, vp, vp], i32),
        "wv_index_build_nsg": ([vp, vp, vp], i32),
        "wv_index_build_vamana": ([vp, vp, vp], i32),
        "wv_index_build_dpg": ([vp, u32, vp], i32),
        "wv_index_build_ssg": ([vp, vp, vp, vp], i32),
        "wv_index_search": ([vp, vp, u32, vp, u64, u32, u32, vp, vp, u32, vp, vp, vp, vp, vp], i32),
        "wv_index_search_dev": ([vp, vp, u32, vp, u64, u32, u32, vp, vp, u32, vp, vp, vp, vp, vp], i32),
        "wv_index_search_dev_noalloc": ([vp, vp, u32, u64, u32, u32, u32, vp, vp, vp, vp, vp], i32),
        "wv_nndescent_build": ([vp, vp, u32, u32, vp, vp, vp, vp, vp, vp], i32),
        "wv_search_greedy": ([vp, vp, vp, u32, vp, u32, vp, u64, u32, u32, vp, vp, u32, vp, vp, vp,
                              vp, vp], i32),
        "wv_exact_knn": ([vp, vp, u32, vp, u32, u32, u32, C.c_int64, vp, vp, vp], i32),
        "wv_gather_rate": ([vp, u32, u32, u32, u32, C.POINTER(C.c_double), C.POINTER(C.c_double)], i32),
        "wv_line_rate": ([vp, C.c_uint64, u32, C.POINTER(C.c_double), C.POINTER(C.c_double)], i32),
        "wv_index_search_ext": ([vp, vp, u32, vp, u64, u32, u32, vp, vp, u32, vp, vp, vp, vp, vp, vp], i32),
        "wv_index_search_ext_dev": ([vp, vp, u32, vp, u64, u32, u32, vp, vp, u32, vp, vp, vp, vp, vp, vp],
                                    i32),
        "wv_index_autotune": ([vp, vp, u32, u64, u32, u32, vp, vp, C.POINTER(C.c_double)], i32),
        "wv_index_autotune_budgets": ([vp, vp, u32, u64, vp, vp, u32, vp, vp, C.POINTER(C.c_double)], i32),
        "wv_search_graph_create": ([vp, vp, u32, u64, u32, u32, vp, vp, u32, vp, vp, vp, vp, vp,
                                    C.POINTER(vp)], i32),
        "wv_search_graph_launch": ([vp, i32], i32),
        "wv_search_graph_destroy": ([vp], i32),
        "wv_mlp_grad": ([vp, vp, vp, vp, u32, C.c_float, vp, C.POINTER(C.c_double)], i32),
        "wv_mlp_train": ([vp, vp, vp, u32, vp, vp, vp], i32),
        "wv_mlp_predict": ([vp, vp, vp, u32, vp], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _p(a: Optional[np.ndarray]):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _c(a, dt):
    return None if a is None else np.ascontiguousarray(a, dtype=dt)


class Context:
    """A device context (one HIP stream, serialised calls)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.wv_ctx_create(device, C.byref(h))
        if rc != WV_OK:
            raise WvError(rc, f"wv_ctx_create(device={device}) failed")
        self.h = h

    def check(self, rc: int):
        if rc != WV_OK:
            raise WvError(rc, self.lib.wv_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.wv_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def exact_knn(self, base, queries, k, self_offset=-1):
        base = _c(base, np.float32)
        queries = _c(queries, np.float32)
        N, d = base.shape
        Q = queries.shape[0]
        ids = np.zeros((Q, k), np.uint32)
        dist = np.zeros((Q, k), np.float64)
        ms = C.c_double(0.0)
        self.check(self.lib.wv_exact_knn(self.h, _p(base), N, _p(queries), Q, d, k, self_offset,
                                         _p(ids), _p(dist), C.byref(ms)))
        return ids, dist, ms.value

    def gather_rate(self, N, d, waves_per_cu, batches=256):
        """Random row gathers of an N x d table scored with the search's
        64-row gather tile, waves_per_cu wavefronts resident a CU; returns
        (GB/s of row bytes, kernel ms)."""
        gbps, ms = C.c_double(0.0), C.c_double(0.0)
        self.check(self.lib.wv_gather_rate(self.h, N, d, waves_per_cu, batches, C.byref(gbps), C.byref(ms)))
        return gbps.value, ms.value

    def line_rate(self, nbytes, iters=64):
        """16-byte loads at random positions of an nbytes device array, 8
        wavefronts a SIMD; returns (G loads/s, kernel ms)."""
        g, ms = C.c_double(0.0), C.c_double(0.0)
        self.check(self.lib.wv_line_rate(self.h, nbytes, iters, C.byref(g), C.byref(ms)))
        return g.value, ms.value

    def mlp_grad(self, model: "TermModel", X, y, pos_weight=1.0):
        X = _c(X, np.float32)
        y = _c(y, np.float32)
        g = np.zeros(321, np.float32)
        loss = C.c_double(0.0)
        self.check(self.lib.wv_mlp_grad(self.h, C.byref(model), _p(X), _p(y), X.shape[0], pos_weight,
                                        _p(g), C.byref(loss)))
        return g, loss.value

    def mlp_train(self, model: "TermModel", X, y, epochs=20, batch=4096, lr=3e-3, pos_weight=1.0,
                  seed=0, beta1=0.9, beta2=0.999, eps=1e-8):
        """Train `model` in place on the device; returns the training statistics."""
        X = _c(X, np.float32)
        y = _c(y, np.float32)
        tp = TrainParams(epochs, batch, lr, beta1, beta2, eps, pos_weight, seed)
        st = TrainStats()
        self.check(self.lib.wv_mlp_train(self.h, _p(X), _p(y), X.shape[0], C.byref(tp), C.byref(model),
                                         C.byref(st)))
        return st.as_dict()

    def mlp_predict(self, model: "TermModel", X):
        X = _c(X, np.float32)
        z = np.zeros(X.shape[0], np.float32)
        self.check(self.lib.wv_mlp_predict(self.h, C.byref(model), _p(X), X.shape[0], _p(z)))
        return z

    def nndescent_build(self, base, K, L, S, R, iters, seed=0, dtype="f32", rerank_mult=0,
                        cand_slots=0, init_ids=None, order=0):
        base = _c(base, np.float32)
        N, d = base.shape
        p = NndParams(K, L, S, R, iters, DTYPES[dtype], rerank_mult, cand_slots, seed, order)
        ids = np.zeros((N, K), np.uint32)
        dist = np.zeros((N, K), np.float32)
        deg = np.zeros(N, np.uint32)
        st = BuildStats()
        ii = _c(init_ids, np.uint32)
        self.check(self.lib.wv_nndescent_build(self.h, _p(base), N, d, C.byref(p), _p(ii), _p(ids),
                                               _p(dist), _p(deg), C.byref(st)))
        return ids, dist, deg, st.as_dict()

    def search_greedy(self, graph, deg, base, queries, L, k, seed=0, entries=None, L_per_q=None,
                      k_per_q=None, k_stride=None):
        graph = _c(graph, np.uint32)
        deg = _c(deg, np.uint32)
        base = _c(base, np.float32)
        queries = _c(queries, np.float32)
        N, d = base.shape
        Q = queries.shape[0]
        if k_stride is None:
            k_stride = int(np.max(k_per_q)) if k_per_q is not None else k
        g = Graph(_p(graph), _p(deg), N, graph.shape[1])
        ids = np.zeros((Q, k_stride), np.uint32)
        dist = np.zeros((Q, k_stride), np.float32)
        ev = np.zeros(Q, np.uint32)
        hp = np.zeros(Q, np.uint32)
        st = SearchStats()
        e, lq, kq = _c(entries, np.uint32), _c(L_per_q, np.uint32), _c(k_per_q, np.uint32)
        self.check(self.lib.wv_search_greedy(self.h, C.byref(g), _p(base), d, _p(queries), Q, _p(e),
                                             seed, L, k, _p(lq), _p(kq), k_stride, _p(ids),
                                             _p(dist), _p(ev), _p(hp), C.byref(st)))
        return ids, dist, ev, hp, st.as_dict()


class SearchGraph:
    """A captured search launch: ``launch()`` replays it on the context's
    stream with the buffers given at capture (rewrite the query rows between
    replays to serve new queries)."""

    def __init__(self, ctx: "Context", h):
        self.ctx, self.h = ctx, h

    def launch(self, sync: bool = True):
        self.ctx.check(self.ctx.lib.wv_search_graph_launch(self.h, 1 if sync else 0))

    def close(self):
        if self.h:
            self.ctx.lib.wv_search_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Index:
    """A device-resident base set and its graph."""

    def __init__(self, ctx: Context, base):
        self.ctx = ctx
        base = _c(base, np.float32)
        self.N, self.d = base.shape
        self.dp = (self.d + 3) & ~3
        h = C.c_void_p()
        ctx.check(ctx.lib.wv_index_create(ctx.h, _p(base), self.N, self.d, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.wv_index_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_graph(self, ids, deg):
        ids = _c(ids, np.uint32)
        deg = _c(deg, np.uint32)
        assert ids.shape[0] == self.N and deg.shape[0] == self.N
        self.ctx.check(self.ctx.lib.wv_index_set_graph(self.h, _p(ids), _p(deg), ids.shape[1]))

    def get_graph(self):
        dm = C.c_uint32(0)
        self.ctx.check(self.ctx.lib.wv_index_get_graph(self.h, None, None, None, C.byref(dm)))
        ids = np.zeros((self.N, dm.value), np.uint32)
        dist = np.zeros((self.N, dm.value), np.float32)
        deg = np.zeros(self.N, np.uint32)
        self.ctx.check(self.ctx.lib.wv_index_get_graph(self.h, _p(ids), _p(dist), _p(deg), None))
        return ids, dist, deg

    def set_search_config(self, waves_per_block=0, row_groups=0, visited_slots=0, merge_buffer=0, schedule=0,
                          waves_per_query=0):
        """Search tuning; schedule 2 serves queries from persistent wavefronts
        through a device work queue (0/1: one wavefront per query);
        waves_per_query 2 or 4 lets that many wavefronts share each query."""
        cfg = SearchConfig(waves_per_block, row_groups, visited_slots, merge_buffer, schedule, waves_per_query)
        self.ctx.check(self.ctx.lib.wv_index_set_search_config(self.h, C.byref(cfg)))

    def build_nndescent(self, K, L, S, R, iters, seed=0, dtype="f32", rerank_mult=0,
                        cand_slots=0, init_ids=None, init_cnt=None, order=0):
        """NN-Descent on the resident base; order: 0 automatic, 1 nodes by id,
        2 grouped by nearest pivot (speed only, the graph is the same)."""
        p = NndParams(K, L, S, R, iters, DTYPES[dtype], rerank_mult, cand_slots, seed, order)
        st = BuildStats()
        ii, ic = _c(init_ids, np.uint32), _c(init_cnt, np.uint32)
        self.ctx.check(self.ctx.lib.wv_index_build_nndescent(self.h, C.byref(p), _p(ii), _p(ic),
                                                             C.byref(st)))
        return st.as_dict()

    def build_nndescent_monitor(self, K, L, S, R, iters, ctrl, truth, seed=0, dtype="f32", rerank_mult=0,
                                cand_slots=0, init_ids=None, init_cnt=None, order=0):
        """NN-Descent with the per-round graph recall of the control nodes
        `ctrl` against their exact neighbour rows `truth`; returns
        (stats, per_round recall)."""
        p = NndParams(K, L, S, R, iters, DTYPES[dtype], rerank_mult, cand_slots, seed, order)
        st = BuildStats()
        ctrl = _c(ctrl, np.uint32)
        truth = _c(truth, np.uint32)
        per = np.zeros(max(iters, 1), np.float64)
        ii, ic = _c(init_ids, np.uint32), _c(init_cnt, np.uint32)
        self.ctx.check(self.ctx.lib.wv_index_build_nndescent_monitor(
            self.h, C.byref(p), _p(ii), _p(ic), _p(ctrl), len(ctrl), _p(truth), truth.shape[1], _p(per),
            C.byref(st)))
        return st.as_dict(), per[:iters]

    def build_efanna(self, K, L, S, R, iters, seed=0, dtype="f32", rerank_mult=0,
                     cand_slots=0, init_ids=None, init_cnt=None, order=0):
        """EFANNA's refine; init_ids: None or an N x stride initial graph."""
        p = NndParams(K, L, S, R, iters, DTYPES[dtype], rerank_mult, cand_slots, seed, order)
        st = BuildStats()
        ii, ic = _c(init_ids, np.uint32), _c(init_cnt, np.uint32)
        stride = 0 if ii is None else ii.shape[1]
        self.ctx.check(self.ctx.lib.wv_index_build_efanna(self.h, C.byref(p), _p(ii), _p(ic), stride,
                                                          C.byref(st)))
        return st.as_dict()

    def build_nsg(self, L, R, C_, ep, seed=0):
        """NSG refine of the index's kNN graph from navigating node ep."""
        p = NsgParams(L, R, C_, ep, seed)
        st = NsgStats()
        self.ctx.check(self.ctx.lib.wv_index_build_nsg(self.h, C.byref(p), C.byref(st)))
        return st.as_dict()

    def build_vamana(self, L, R, C_, ep, alpha, seed=0):
        """One Vamana pass over the index's graph from navigating node ep."""
        p = VamanaParams(L, R, C_, ep, alpha, 0, seed)
        st = NsgStats()
        self.ctx.check(self.ctx.lib.wv_index_build_vamana(self.h, C.byref(p), C.byref(st)))
        return st.as_dict()

    def build_ssg(self, L, R, angle=60.0, n_try=10, seed=0):
        """SSG refine of the index's kNN graph; returns (stats, eps)."""
        p = SsgParams(L, R, angle, n_try, seed)
        st = NsgStats()
        eps = np.zeros(n_try, np.uint32)
        self.ctx.check(self.ctx.lib.wv_index_build_ssg(self.h, C.byref(p), _p(eps), C.byref(st)))
        return st.as_dict(), eps

    def build_dpg(self, range_):
        """DPG refine of the index's kNN graph (`range_` a node, then mirrored)."""
        st = NsgStats()
        self.ctx.check(self.ctx.lib.wv_index_build_dpg(self.h, range_, C.byref(st)))
        return st.as_dict()

    def search(self, queries, L, k, seed=0, entries=None, L_per_q=None, k_per_q=None,
               k_stride=None):
        queries = _c(queries, np.float32)
        Q = queries.shape[0]
        if k_stride is None:
            k_stride = int(np.max(k_per_q)) if k_per_q is not None else k
        ids = np.zeros((Q, k_stride), np.uint32)
        dist = np.zeros((Q, k_stride), np.float32)
        ev = np.zeros(Q, np.uint32)
        hp = np.zeros(Q, np.uint32)
        st = SearchStats()
        e, lq, kq = _c(entries, np.uint32), _c(L_per_q, np.uint32), _c(k_per_q, np.uint32)
        self.ctx.check(self.ctx.lib.wv_index_search(self.h, _p(queries), Q, _p(e), seed, L, k,
                                                    _p(lq), _p(kq), k_stride, _p(ids), _p(dist),
                                                    _p(ev), _p(hp), C.byref(st)))
        return ids, dist, ev, hp, st.as_dict()

    def search_ext(self, queries, L, k, seed=0, model: Optional["TermModel"] = None, trace_cap=0,
                   want_settle=False, L_per_q=None, k_per_q=None, k_stride=None, entries=None):
        """Search with the learned stop rule and/or per-hop feature traces.

        Returns ids, dist, evals, hops, stats, trace (Q x cap x 8 or None),
        trace_len, settle_hop (or None)."""
        queries = _c(queries, np.float32)
        Q = queries.shape[0]
        if k_stride is None:
            k_stride = int(np.max(k_per_q)) if k_per_q is not None else k
        ids = np.zeros((Q, k_stride), np.uint32)
        dist = np.zeros((Q, k_stride), np.float32)
        ev = np.zeros(Q, np.uint32)
        hp = np.zeros(Q, np.uint32)
        trace = np.zeros((Q, trace_cap, 8), np.float32) if trace_cap else None
        tlen = np.zeros(Q, np.uint32) if trace_cap else None
        settle = np.zeros(Q, np.uint32) if (want_settle or trace_cap) else None
        ext = SearchExt(C.addressof(model) if model is not None else None, _p(trace) and _p(trace).value,
                        _p(tlen) and _p(tlen).value, _p(settle) and _p(settle).value, trace_cap)
        st = SearchStats()
        e, lq, kq = _c(entries, np.uint32), _c(L_per_q, np.uint32), _c(k_per_q, np.uint32)
        self.ctx.check(self.ctx.lib.wv_index_search_ext(self.h, _p(queries), Q, _p(e), seed, L, k, _p(lq),
                                                        _p(kq), k_stride, _p(ids), _p(dist), _p(ev), _p(hp),
                                                        C.byref(ext), C.byref(st)))
        return ids, dist, ev, hp, st.as_dict(), trace, tlen, settle

    def search_ext_dev(self, d_queries: int, Q: int, L: int, k: int, d_out_ids: int,
                       model: Optional["TermModel"] = None, d_out_dist: int = 0, seed: int = 0,
                       d_L: int = 0, d_k: int = 0, k_stride: Optional[int] = None, d_evals: int = 0,
                       d_hops: int = 0, want_stats: bool = True):
        """search_dev with the learned stop rule (model is a host struct)."""
        st = SearchStats()
        vp = lambda x: C.c_void_p(x) if x else None
        ext = SearchExt(C.addressof(model) if model is not None else None, None, None, None, 0)
        self.ctx.check(self.ctx.lib.wv_index_search_ext_dev(
            self.h, vp(d_queries), Q, None, seed, L, k, vp(d_L), vp(d_k),
            k_stride if k_stride is not None else k, vp(d_out_ids), vp(d_out_dist),
            vp(d_evals), vp(d_hops), C.byref(ext), C.byref(st) if want_stats else None))
        return st.as_dict()

    def autotune_dev(self, d_queries: int, Q: int, L: int, k: int, seed: int = 0,
                     model: Optional["TermModel"] = None):
        """Pick the fastest search tuning for (L, k[, model]) on Q sample
        queries (device rows); the index keeps it.  Returns (config dict,
        kernel ms of the sample)."""
        cfg = SearchConfig()
        ms = C.c_double(0.0)
        ext = SearchExt(C.addressof(model), None, None, None, 0) if model is not None else None
        self.ctx.check(self.ctx.lib.wv_index_autotune(self.h, C.c_void_p(d_queries), Q, seed, L, k,
                                                      C.byref(ext) if ext is not None else None,
                                                      C.byref(cfg), C.byref(ms)))
        return {f: getattr(cfg, f) for f, _ in cfg._fields_}, ms.value

    def autotune_budgets_dev(self, d_queries: int, Q: int, d_L: int, d_k: int, k_stride: int, seed: int = 0,
                             model: Optional["TermModel"] = None):
        """autotune_dev under the sample queries' own per-query L and k
        (device arrays of Q values each); the index keeps the winner."""
        cfg = SearchConfig()
        ms = C.c_double(0.0)
        ext = SearchExt(C.addressof(model), None, None, None, 0) if model is not None else None
        self.ctx.check(self.ctx.lib.wv_index_autotune_budgets(
            self.h, C.c_void_p(d_queries), Q, seed, C.c_void_p(d_L), C.c_void_p(d_k), k_stride,
            C.byref(ext) if ext is not None else None, C.byref(cfg), C.byref(ms)))
        return {f: getattr(cfg, f) for f, _ in cfg._fields_}, ms.value

    def search_graph(self, d_queries: int, Q: int, L: int, k: int, d_out_ids: int, seed: int = 0,
                     d_out_dist: int = 0, d_L: int = 0, d_k: int = 0, k_stride: Optional[int] = None,
                     d_evals: int = 0, d_hops: int = 0, model: Optional["TermModel"] = None) -> "SearchGraph":
        """Capture the device-buffer search once as a hipGraph (see SearchGraph)."""
        vp = lambda x: C.c_void_p(x) if x else None
        ext = SearchExt(C.addressof(model), None, None, None, 0) if model is not None else None
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.wv_search_graph_create(
            self.h, vp(d_queries), Q, seed, L, k, vp(d_L), vp(d_k), k_stride if k_stride is not None else k,
            vp(d_out_ids), vp(d_out_dist), vp(d_evals), vp(d_hops), C.byref(ext) if ext is not None else None,
            C.byref(h)))
        return SearchGraph(self.ctx, h)

    def autotune(self, queries, L: int, k: int, seed: int = 0, model: Optional["TermModel"] = None):
        """autotune_dev on host sample queries (copied to the device)."""
        from .devmem import DeviceArray, pad_rows
        dq = DeviceArray.from_host(pad_rows(_c(queries, np.float32)))
        try:
            return self.autotune_dev(dq.ptr, len(queries), L, k, seed=seed, model=model)
        finally:
            dq.free()

    def search_dev(self, d_queries: int, Q: int, L: int, k: int, d_out_ids: int,
                   d_out_dist: int = 0, seed: int = 0, d_entries: int = 0, d_L: int = 0,
                   d_k: int = 0, k_stride: Optional[int] = None, d_evals: int = 0,
                   d_hops: int = 0, want_stats: bool = True):
        """Search with device pointers (ints, e.g. torch ``data_ptr()``)."""
        st = SearchStats()
        vp = lambda x: C.c_void_p(x) if x else None
        self.ctx.check(self.ctx.lib.wv_index_search_dev(
            self.h, vp(d_queries), Q, vp(d_entries), seed, L, k, vp(d_L), vp(d_k),
            k_stride if k_stride is not None else k, vp(d_out_ids), vp(d_out_dist),
            vp(d_evals), vp(d_hops), C.byref(st) if want_stats else None))
        return st.as_dict()

    def search_dev_noalloc(self, d_queries: int, Q: int, L: int, k: int, d_out_ids: int,
                           d_out_dist: int = 0, seed: int = 0, k_stride: Optional[int] = None,
                           d_evals: int = 0, d_hops: int = 0, want_stats: bool = True):
        """wv_index_search_dev_noalloc: random entries, one L and k, no device
        allocation inside the call (its scratch is a fixed array of the
        library, WV_NOALLOC_MAX_Q queries a call)."""
        st = SearchStats()
        vp = lambda x: C.c_void_p(x) if x else None
        self.ctx.check(self.ctx.lib.wv_index_search_dev_noalloc(
            self.h, vp(d_queries), Q, seed, L, k, k_stride if k_stride is not None else k,
            vp(d_out_ids), vp(d_out_dist), vp(d_evals), vp(d_hops), C.byref(st) if want_stats else None))
        return st.as_dict()
