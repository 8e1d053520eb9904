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
]


class WvError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"weavess_hip error {code}: {msg}")
        self.code = code


class NndParams(C.Structure):
    _fields_ = [("K", C.c_uint32), ("L", C.c_uint32), ("S", C.c_uint32), ("R", C.c_uint32),
                ("iters", C.c_uint32), ("dtype", C.c_uint32), ("rerank_mult", C.c_uint32),
                ("cand_slots", C.c_uint32), ("seed", C.c_uint64)]


class BuildStats(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("init_ms", C.c_double), ("sample_ms", C.c_double),
                ("join_ms", C.c_double), ("merge_ms", C.c_double), ("finalize_ms", C.c_double),
                ("pairs", C.c_uint64), ("proposals", C.c_uint64), ("rows_gathered", C.c_uint64),
                ("iters_run", C.c_uint32), ("cand_slots", C.c_uint32)]

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
                ("visited_slots", C.c_uint32)]


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
        "wv_index_search": ([vp, vp, u32, vp, u64, u32, u32, vp, vp, u32, vp, vp, vp, vp, vp], i32),
        "wv_index_search_dev": ([vp, vp, u32, vp, u64, u32, u32, vp, vp, u32, vp, vp, vp, vp, vp], i32),
        "wv_nndescent_build": ([vp, vp, u32, u32, vp, vp, vp, vp, vp, vp], i32),
        "wv_search_greedy": ([vp, vp, vp, u32, vp, u32, vp, u64, u32, u32, vp, vp, u32, vp, vp, vp,
                              vp, vp], i32),
        "wv_exact_knn": ([vp, vp, u32, vp, u32, u32, u32, C.c_int64, vp, vp, vp], i32),
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

    def nndescent_build(self, base, K, L, S, R, iters, seed=0, dtype="f32", rerank_mult=0,
                        cand_slots=0, init_ids=None):
        base = _c(base, np.float32)
        N, d = base.shape
        p = NndParams(K, L, S, R, iters, DTYPES[dtype], rerank_mult, cand_slots, seed)
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

    def set_search_config(self, waves_per_block=0, row_groups=0, visited_slots=0):
        cfg = SearchConfig(waves_per_block, row_groups, visited_slots)
        self.ctx.check(self.ctx.lib.wv_index_set_search_config(self.h, C.byref(cfg)))

    def build_nndescent(self, K, L, S, R, iters, seed=0, dtype="f32", rerank_mult=0,
                        cand_slots=0, init_ids=None, init_cnt=None):
        p = NndParams(K, L, S, R, iters, DTYPES[dtype], rerank_mult, cand_slots, seed)
        st = BuildStats()
        ii, ic = _c(init_ids, np.uint32), _c(init_cnt, np.uint32)
        self.ctx.check(self.ctx.lib.wv_index_build_nndescent(self.h, C.byref(p), _p(ii), _p(ic),
                                                             C.byref(st)))
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
