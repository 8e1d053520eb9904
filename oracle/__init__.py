"""ctypes front of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py may import this package; the product path never does.  It loads
``oracle/libwv_oracle.so`` (built by ``make -C oracle``) and, for pinning the
restatement against the reference's own headers, ``oracle/_ref/libwv_ref.so``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwv_oracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libwv_ref.so")

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class NndParams(C.Structure):
    _fields_ = [("K", C.c_uint32), ("L", C.c_uint32), ("S", C.c_uint32),
                ("R", C.c_uint32), ("iters", C.c_uint32),
                ("cand_slots", C.c_uint32), ("sweep", C.c_uint32),
                ("reserved", C.c_uint32), ("seed", C.c_uint64)]


class TermModel(C.Structure):
    """Same layout as wv_term_model / wvo_term_model."""
    _fields_ = [("w1", (C.c_float * 8) * 32), ("b1", C.c_float * 32), ("w2", C.c_float * 32),
                ("b2", C.c_float), ("mean", C.c_float * 8), ("inv_std", C.c_float * 8),
                ("logit_threshold", C.c_float), ("min_hops", C.c_uint32)]


class NndIterStats(C.Structure):
    _fields_ = [("pairs", C.c_uint64), ("inserts", C.c_uint64)]


def build(ref: bool = True) -> None:
    """Compile the oracle (and the reference shim when the tree is present)."""
    targets = ["all"]
    if ref and os.path.isdir("/root/reference/include/weavess"):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(ref=False)
        L = C.CDLL(LIB_PATH)
        L.wvo_dist_ref.restype = C.c_float
        L.wvo_dist_ref.argtypes = [_f32p, _f32p, C.c_uint32]
        L.wvo_dist_canon.restype = C.c_float
        L.wvo_dist_canon.argtypes = [_f32p, _f32p, C.c_uint32]
        L.wvo_dist_f64.restype = C.c_double
        L.wvo_dist_f64.argtypes = [_f32p, _f32p, C.c_uint32]
        L.wvo_stream.restype = C.c_uint64
        L.wvo_stream.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.wvo_draw.restype = C.c_uint32
        L.wvo_draw.argtypes = [C.c_uint64, C.c_uint64]
        L.wvo_gen_random.restype = C.c_int
        L.wvo_gen_random.argtypes = [C.c_uint64, _u32p, C.c_uint32, C.c_uint32]
        L.wvo_gen_random_draws.restype = C.c_int
        L.wvo_gen_random_draws.argtypes = [_u32p, _u32p, C.c_uint32, C.c_uint32]
        L.wvo_init_random.restype = C.c_int
        L.wvo_init_random.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                                      _u32p, _f32p, _u32p, C.c_int]
        L.wvo_nndescent.restype = C.c_int
        L.wvo_nndescent.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.POINTER(NndParams),
                                    C.c_void_p, C.c_void_p, C.c_void_p,
                                    _u32p, _f32p, _u32p, C.c_void_p, C.c_int]
        L.wvo_nndescent_state.restype = C.c_int
        L.wvo_nndescent_state.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.POINTER(NndParams),
                                          C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_uint32,
                                          C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_int]
        L.wvo_nndescent_efanna.restype = C.c_int
        L.wvo_nndescent_efanna.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.POINTER(NndParams),
                                           C.c_void_p, C.c_void_p, C.c_uint32,
                                           _u32p, _f32p, _u32p, C.c_void_p, C.c_int]
        for fn in ("wvo_nsg", "wvo_nsg_conn"):
            getattr(L, fn).restype = C.c_int
        L.wvo_nsg.argtypes = [_f32p, C.c_uint32, C.c_uint32, _u32p, _u32p, C.c_uint32, C.c_uint32,
                              C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, _u32p, _f32p, _u32p,
                              C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_int]
        L.wvo_nsg_conn.argtypes = [_f32p, C.c_uint32, C.c_uint32, _u32p, _u32p, C.c_uint32, C.c_uint32,
                                   C.c_uint32, C.c_uint32, C.c_uint64, _u32p, _f32p, _u32p, C.c_uint32,
                                   C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_int]
        L.wvo_ssg_threshold.restype = C.c_float
        L.wvo_ssg_threshold.argtypes = [C.c_float]
        L.wvo_ssg.restype = C.c_int
        L.wvo_ssg.argtypes = [_f32p, C.c_uint32, C.c_uint32, _u32p, _u32p, C.c_uint32, C.c_uint32, C.c_uint32,
                              C.c_float, C.c_uint32, C.c_uint64, _u32p, _f32p, _u32p, C.c_uint32,
                              C.POINTER(C.c_uint32), _u32p, C.POINTER(C.c_uint32), C.c_int]
        L.wvo_dpg.restype = C.c_int
        L.wvo_dpg.argtypes = [_f32p, C.c_uint32, C.c_uint32, _u32p, _u32p, C.c_uint32, C.c_uint32, _u32p,
                              _f32p, _u32p, C.c_uint32, C.POINTER(C.c_uint32), C.c_int]
        L.wvo_vamana.restype = C.c_int
        L.wvo_vamana.argtypes = [_f32p, C.c_uint32, C.c_uint32, _u32p, _u32p, C.c_uint32, C.c_uint32,
                                 C.c_uint32, C.c_uint32, C.c_float, C.c_uint32, C.c_uint64, _u32p, _f32p,
                                 _u32p, C.c_uint32, C.POINTER(C.c_uint32), C.c_int]
        L.wvo_search_greedy.restype = C.c_int
        L.wvo_search_greedy.argtypes = [_u32p, _u32p, C.c_uint32, _f32p, C.c_uint32, C.c_uint32,
                                        _f32p, C.c_uint32, C.c_void_p, C.c_uint64,
                                        C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                        C.c_uint32, _u32p, _f32p, _u32p, _u32p, C.c_int]
        L.wvo_search_greedy_ext.restype = C.c_int
        L.wvo_search_greedy_ext.argtypes = [_u32p, _u32p, C.c_uint32, _f32p, C.c_uint32, C.c_uint32,
                                            _f32p, C.c_uint32, C.c_void_p, C.c_uint64,
                                            C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                            C.c_uint32, _u32p, _f32p, _u32p, _u32p, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int]
        L.wvo_term_features.restype = None
        L.wvo_term_features.argtypes = [C.c_float] * 8 + [_f32p]
        L.wvo_mlp_grad.restype = C.c_int
        L.wvo_mlp_grad.argtypes = [C.c_void_p, _f32p, _f32p, C.c_uint32, C.c_float, _f64p,
                                   C.POINTER(C.c_double)]
        L.wvo_mlp_logit.restype = C.c_int
        L.wvo_mlp_logit.argtypes = [C.c_void_p, _f32p, C.c_uint32, C.c_int, _f64p]
        L.wvo_exact_knn.restype = C.c_int
        L.wvo_exact_knn.argtypes = [_f32p, C.c_uint32, _f32p, C.c_uint32, C.c_uint32,
                                    C.c_uint32, C.c_int64, _u32p, _f64p, C.c_int]
        L.wvo_recall.restype = C.c_double
        L.wvo_recall.argtypes = [_u32p, C.c_uint32, C.c_uint32, C.c_uint32, _u32p, C.c_uint32]
        L.wvo_graph_quality.restype = C.c_double
        L.wvo_graph_quality.argtypes = [_u32p, _u32p, C.c_uint32, C.c_void_p, C.c_uint32,
                                        _u32p, C.c_uint32, C.c_uint32]
        _lib = L
    return _lib


def ref_lib():
    """The reference-header shim, or None where it was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB_PATH):
        R = C.CDLL(REF_LIB_PATH)
        R.wvr_dist.restype = C.c_float
        R.wvr_dist.argtypes = [_f32p, _f32p, C.c_uint32]
        R.wvr_gen_random.restype = None
        R.wvr_gen_random.argtypes = [C.c_uint32, _u32p, C.c_uint32, C.c_uint32]
        R.wvr_mt19937.restype = None
        R.wvr_mt19937.argtypes = [C.c_uint32, _u32p, C.c_uint32]
        _ref = R
    return _ref


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------- distances
def dist_ref(a, b) -> float:
    a, b = _c(a, np.float32), _c(b, np.float32)
    return float(lib().wvo_dist_ref(a, b, a.size))


def dist_canon(a, b) -> float:
    a, b = _c(a, np.float32), _c(b, np.float32)
    return float(lib().wvo_dist_canon(a, b, a.size))


def dist_f64(a, b) -> float:
    a, b = _c(a, np.float32), _c(b, np.float32)
    return float(lib().wvo_dist_f64(a, b, a.size))


# ---------------------------------------------------------------- rng
TAG_INIT = 0x494E4954
TAG_ENTRY = 0x454E5452


def stream(seed: int, tag: int, idx: int) -> int:
    return int(lib().wvo_stream(seed, tag, idx))


def draw(s: int, k: int) -> int:
    return int(lib().wvo_draw(s, k))


def gen_random(s: int, size: int, N: int) -> np.ndarray:
    out = np.zeros(size, np.uint32)
    if lib().wvo_gen_random(s, out, size, N) != 0:
        raise ValueError("gen_random: size must be < N")
    return out


def gen_random_draws(draws, size: int, N: int) -> np.ndarray:
    out = np.zeros(size, np.uint32)
    d = _c(draws, np.uint32)
    assert d.size >= size + 1
    if lib().wvo_gen_random_draws(d, out, size, N) != 0:
        raise ValueError("gen_random_draws: size must be < N")
    return out


# ---------------------------------------------------------------- nn-descent
def init_random(base, S: int, seed: int, nthreads: int = 0):
    base = _c(base, np.float32)
    N, d = base.shape
    ids = np.zeros((N, S), np.uint32)
    dist = np.zeros((N, S), np.float32)
    cnt = np.zeros(N, np.uint32)
    if lib().wvo_init_random(base, N, d, S, seed, ids, dist, cnt, nthreads) != 0:
        raise ValueError("init_random: bad arguments")
    return ids, dist, cnt


def nndescent(base, K, L, S, R, iters, seed=0, cand_slots=0, init=None,
              nthreads: int = 0, with_stats: bool = False, sweep: bool = False):
    """InitRandom + NN-Descent + PruneNaive.  cand_slots=0: the reference's
    exact bounded insertion; > 0: the slot table the device uses.  sweep:
    the reference's single-thread reverse test (a neighbour visited earlier
    in the sweep reads a heap leaf as its worst) instead of the two-phase
    one."""
    base = _c(base, np.float32)
    N, d = base.shape
    p = NndParams(K, L, S, R, iters, cand_slots, int(sweep), 0, seed)
    ids = np.zeros((N, K), np.uint32)
    dist = np.zeros((N, K), np.float32)
    deg = np.zeros(N, np.uint32)
    st = (NndIterStats * max(iters, 1))()
    if init is not None:
        i_ids, i_dist, i_cnt = (_c(init[0], np.uint32), _c(init[1], np.float32),
                                _c(init[2], np.uint32))
        args = (_ptr(i_ids), _ptr(i_dist), _ptr(i_cnt))
    else:
        args = (None, None, None)
    rc = lib().wvo_nndescent(base, N, d, C.byref(p), *args, ids, dist, deg,
                             C.cast(st, C.c_void_p), nthreads)
    if rc != 0:
        raise ValueError("nndescent: bad arguments")
    if with_stats:
        return ids, dist, deg, [(s.pairs, s.inserts) for s in st[:iters]]
    return ids, dist, deg


def nndescent_state(base, K, L, S, R, iters, seed=0, cand_slots=0, init=None,
                    nthreads: int = 1, new_cap=None, old_cap=None, sweep: bool = False):
    """State after ``iters`` (join, update) rounds, before the final sort:
    dict of pool ids/dist/flag (N x L, ``pool_cnt`` valid), ``M``, the sample
    lists ``nn_new``/``nn_old`` (Python lists per node) and per-round
    (pairs, inserts).  Single-threaded by default so that exact insertion is
    deterministic."""
    base = _c(base, np.float32)
    N, d = base.shape
    p = NndParams(K, L, S, R, iters, cand_slots, int(sweep), 0, seed)
    new_cap = new_cap or max(2 * S, S + R)
    old_cap = old_cap or max(L, 2 * R)
    pid = np.zeros((N, L), np.uint32)
    pdist = np.zeros((N, L), np.float32)
    pflag = np.zeros((N, L), np.uint8)
    pcnt = np.zeros(N, np.uint32)
    M = np.zeros(N, np.uint32)
    nn_new = np.zeros((N, new_cap), np.uint32)
    nn_old = np.zeros((N, old_cap), np.uint32)
    ncnt = np.zeros(N, np.uint32)
    ocnt = np.zeros(N, np.uint32)
    st = (NndIterStats * max(iters, 1))()
    if init is not None:
        i_ids, i_dist, i_cnt = (_c(init[0], np.uint32), _c(init[1], np.float32),
                                _c(init[2], np.uint32))
        args = (_ptr(i_ids), _ptr(i_dist), _ptr(i_cnt))
    else:
        args = (None, None, None)
    rc = lib().wvo_nndescent_state(base, N, d, C.byref(p), *args, _ptr(pid), _ptr(pdist),
                                   _ptr(pflag), _ptr(pcnt), _ptr(M), _ptr(nn_new), _ptr(ncnt),
                                   new_cap, _ptr(nn_old), _ptr(ocnt), old_cap,
                                   C.cast(st, C.c_void_p), nthreads)
    if rc != 0:
        raise ValueError("nndescent_state: bad arguments (%d)" % rc)
    return {
        "pool_ids": pid, "pool_dist": pdist, "pool_flag": pflag, "pool_cnt": pcnt, "M": M,
        "nn_new": [list(nn_new[i, :ncnt[i]]) for i in range(N)],
        "nn_old": [list(nn_old[i, :ocnt[i]]) for i in range(N)],
        "stats": [(s.pairs, s.inserts) for s in st[:iters]],
    }


def nndescent_efanna(base, K, L, S, R, iters, seed=0, cand_slots=0, init=None,
                     nthreads: int = 0, with_stats: bool = False):
    """EFANNA's refine.  init: None (InitRandom's graph) or (ids N x stride,
    cnt[N] or None) of any initial graph with stride <= L."""
    base = _c(base, np.float32)
    N, d = base.shape
    p = NndParams(K, L, S, R, iters, cand_slots, 0, 0, seed)
    ids = np.zeros((N, K), np.uint32)
    dist = np.zeros((N, K), np.float32)
    deg = np.zeros(N, np.uint32)
    st = (NndIterStats * max(iters, 1))()
    if init is not None:
        i_ids = _c(init[0], np.uint32)
        i_cnt = None if init[1] is None else _c(init[1], np.uint32)
        args = (_ptr(i_ids), _ptr(i_cnt), i_ids.shape[1])
    else:
        args = (None, None, 0)
    rc = lib().wvo_nndescent_efanna(base, N, d, C.byref(p), *args, ids, dist, deg,
                                    C.cast(st, C.c_void_p), nthreads)
    if rc != 0:
        raise ValueError("nndescent_efanna: bad arguments")
    if with_stats:
        return ids, dist, deg, [(s.pairs, s.inserts) for s in st[:iters]]
    return ids, dist, deg


def _graph_call(fn, N, first_width, *args):
    wcap = first_width
    while True:
        ids = np.zeros((N, wcap), np.uint32)
        dist = np.zeros((N, wcap), np.float32)
        deg = np.zeros(N, np.uint32)
        w, e = C.c_uint32(0), C.c_uint32(0)
        rc = fn(*args, ids, dist, deg, wcap, C.byref(w), C.byref(e), 0)
        if rc == -2:
            wcap = w.value
            continue
        if rc != 0:
            raise ValueError(f"{fn.__name__}: bad arguments")
        return ids[:, :w.value], dist[:, :w.value], deg, e.value


def nsg(base, knn_ids, knn_deg, L, R, C_, ep, seed=0):
    """NSG refine of a kNN graph; returns (ids, dist, deg, conn_edges)."""
    base = _c(base, np.float32)
    knn_ids = _c(knn_ids, np.uint32)
    knn_deg = _c(knn_deg, np.uint32)
    N, d = base.shape
    return _graph_call(lib().wvo_nsg, N, R + 8, base, N, d, knn_ids, knn_deg, knn_ids.shape[1], L, R, C_, ep,
                       seed)


def vamana(base, ids, deg, L, R, C_, alpha, ep, seed=0):
    """One Vamana pass over graph ids/deg; returns (ids, dist, deg)."""
    base = _c(base, np.float32)
    ids = _c(ids, np.uint32)
    deg = _c(deg, np.uint32)
    N, d = base.shape
    wcap = R
    while True:
        oi = np.zeros((N, wcap), np.uint32)
        od = np.zeros((N, wcap), np.float32)
        odeg = np.zeros(N, np.uint32)
        w = C.c_uint32(0)
        rc = lib().wvo_vamana(base, N, d, ids, deg, ids.shape[1], L, R, C_, alpha, ep, seed, oi, od, odeg, wcap,
                              C.byref(w), 0)
        if rc == -2:
            wcap = w.value
            continue
        if rc != 0:
            raise ValueError("vamana: bad arguments")
        return oi[:, :w.value], od[:, :w.value], odeg


def ssg(base, knn_ids, knn_deg, L, R, angle=60.0, n_try=10, seed=0):
    """SSG refine; returns (ids, dist, deg, eps, conn_edges)."""
    base = _c(base, np.float32)
    knn_ids = _c(knn_ids, np.uint32)
    knn_deg = _c(knn_deg, np.uint32)
    N, d = base.shape
    wcap = R + 4
    eps = np.zeros(n_try, np.uint32)
    while True:
        oi = np.zeros((N, wcap), np.uint32)
        od = np.zeros((N, wcap), np.float32)
        odeg = np.zeros(N, np.uint32)
        w, e = C.c_uint32(0), C.c_uint32(0)
        rc = lib().wvo_ssg(base, N, d, knn_ids, knn_deg, knn_ids.shape[1], L, R, angle, n_try, seed, oi, od,
                           odeg, wcap, C.byref(w), eps, C.byref(e), 0)
        if rc == -2:
            wcap = w.value
            continue
        if rc != 0:
            raise ValueError("ssg: bad arguments")
        return oi[:, :w.value], od[:, :w.value], odeg, eps, e.value


def ssg_threshold(angle):
    return float(lib().wvo_ssg_threshold(angle))


def dpg(base, ids, deg, range_):
    """DPG's prune of graph ids/deg to `range_` a node, then the reverse
    union; returns (ids, dist, deg)."""
    base = _c(base, np.float32)
    ids = _c(ids, np.uint32)
    deg = _c(deg, np.uint32)
    N, d = base.shape
    wcap = 2 * range_
    while True:
        oi = np.zeros((N, wcap), np.uint32)
        od = np.zeros((N, wcap), np.float32)
        odeg = np.zeros(N, np.uint32)
        w = C.c_uint32(0)
        rc = lib().wvo_dpg(base, N, d, ids, deg, ids.shape[1], range_, oi, od, odeg, wcap, C.byref(w), 0)
        if rc == -2:
            wcap = w.value
            continue
        if rc != 0:
            raise ValueError("dpg: bad arguments")
        return oi[:, :w.value], od[:, :w.value], odeg


def nsg_conn(base, ids, deg, L, C_, ep, seed=0):
    """NSG's connectivity pass alone; returns (ids, dist, deg, edges)."""
    base = _c(base, np.float32)
    ids = _c(ids, np.uint32)
    deg = _c(deg, np.uint32)
    N, d = base.shape
    return _graph_call(lib().wvo_nsg_conn, N, ids.shape[1] + 8, base, N, d, ids, deg, ids.shape[1], L, C_, ep,
                       seed)


# ---------------------------------------------------------------- search
def search_greedy(graph, deg, base, queries, L, k, seed=0, entries=None,
                  L_per_q=None, k_per_q=None, k_stride=None, nthreads: int = 0):
    graph = _c(graph, np.uint32)
    deg = _c(deg, np.uint32)
    base = _c(base, np.float32)
    queries = _c(queries, np.float32)
    N, d = base.shape
    Q = queries.shape[0]
    deg_max = graph.shape[1]
    if k_stride is None:
        k_stride = int(np.max(k_per_q)) if k_per_q is not None else k
    out_ids = np.zeros((Q, k_stride), np.uint32)
    out_dist = np.zeros((Q, k_stride), np.float32)
    evals = np.zeros(Q, np.uint32)
    hops = np.zeros(Q, np.uint32)
    ent = None if entries is None else _c(entries, np.uint32)
    lq = None if L_per_q is None else _c(L_per_q, np.uint32)
    kq = None if k_per_q is None else _c(k_per_q, np.uint32)
    rc = lib().wvo_search_greedy(graph, deg, deg_max, base, N, d, queries, Q, _ptr(ent), seed,
                                 L, k, _ptr(lq), _ptr(kq), k_stride, out_ids, out_dist,
                                 evals, hops, nthreads)
    if rc != 0:
        raise ValueError("search_greedy: bad arguments")
    return out_ids, out_dist, evals, hops


# ---------------------------------------------------------------- exact / metrics
def exact_knn(base, queries, k, self_offset=-1, nthreads: int = 0):
    base = _c(base, np.float32)
    queries = _c(queries, np.float32)
    N, d = base.shape
    Q = queries.shape[0]
    ids = np.zeros((Q, k), np.uint32)
    dist = np.zeros((Q, k), np.float64)
    if lib().wvo_exact_knn(base, N, queries, Q, d, k, self_offset, ids, dist, nthreads) != 0:
        raise ValueError("exact_knn: bad arguments")
    return ids, dist


def recall(res, gt, k) -> float:
    res = _c(res, np.uint32)
    gt = _c(gt, np.uint32)
    return float(lib().wvo_recall(res, res.shape[0], k, res.shape[1], gt, gt.shape[1]))


def graph_quality(ids, deg, gt, rows=None, control=10) -> float:
    ids = _c(ids, np.uint32)
    deg = _c(deg, np.uint32)
    gt = _c(gt, np.uint32)
    r = None if rows is None else _c(rows, np.uint32)
    return float(lib().wvo_graph_quality(ids, deg, ids.shape[1], _ptr(r), gt.shape[0], gt,
                                         gt.shape[1], control))


# ---------------------------------------------------------------- learned termination
def search_greedy_ext(graph, deg, base, queries, L, k, seed=0, entries=None, model=None,
                      trace_cap=0, L_per_q=None, k_per_q=None, nthreads: int = 0):
    """Greedy search with per-hop feature traces and/or the learned stop."""
    graph = _c(graph, np.uint32)
    deg = _c(deg, np.uint32)
    base = _c(base, np.float32)
    queries = _c(queries, np.float32)
    N, d = base.shape
    Q = queries.shape[0]
    k_stride = int(np.max(k_per_q)) if k_per_q is not None else k
    out_ids = np.zeros((Q, k_stride), np.uint32)
    out_dist = np.zeros((Q, k_stride), np.float32)
    evals = np.zeros(Q, np.uint32)
    hops = np.zeros(Q, np.uint32)
    ent = None if entries is None else _c(entries, np.uint32)
    lq = None if L_per_q is None else _c(L_per_q, np.uint32)
    kq = None if k_per_q is None else _c(k_per_q, np.uint32)
    trace = np.zeros((Q, max(trace_cap, 1), 8), np.float32) if trace_cap else None
    tlen = np.zeros(Q, np.uint32)
    settle = np.zeros(Q, np.uint32)
    rc = lib().wvo_search_greedy_ext(graph, deg, graph.shape[1], base, N, d, queries, Q, _ptr(ent), seed,
                                     L, k, _ptr(lq), _ptr(kq), k_stride, out_ids, out_dist, evals, hops,
                                     None if model is None else C.cast(C.byref(model), C.c_void_p),
                                     _ptr(trace), _ptr(tlen), _ptr(settle), trace_cap, nthreads)
    if rc != 0:
        raise ValueError("search_greedy_ext: bad arguments")
    return out_ids, out_dist, evals, hops, trace, tlen, settle


def term_features(hops, kth, best, entry, kth4, stale, L, k):
    f = np.zeros(8, np.float32)
    lib().wvo_term_features(hops, kth, best, entry, kth4, stale, L, k, f)
    return f


def mlp_grad(model, X, y, pos_weight=1.0):
    X = _c(X, np.float32)
    y = _c(y, np.float32)
    g = np.zeros(321, np.float64)
    loss = C.c_double(0.0)
    if lib().wvo_mlp_grad(C.cast(C.byref(model), C.c_void_p), X, y, X.shape[0], pos_weight, g,
                          C.byref(loss)) != 0:
        raise ValueError("mlp_grad: bad arguments")
    return g, loss.value


def mlp_logit(model, X, bf16=True):
    X = _c(X, np.float32)
    out = np.zeros(X.shape[0], np.float64)
    lib().wvo_mlp_logit(C.cast(C.byref(model), C.c_void_p), X, X.shape[0], 1 if bf16 else 0, out)
    return out
