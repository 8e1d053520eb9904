"""Pinning the CPU oracle (CPU only).

The reference ships no tests, golden vectors or fixtures for this path, so
the oracle is pinned by: the reference's own headers compiled in place
(oracle/_ref: distance.h, util.h), closed forms, brute force on small
inputs, special cases that reduce to exact kNN, and hand-worked known
answers whose expected values are written out below.
"""
import os

import numpy as np
import pytest

import oracle as O
from tests.conftest import chain_points

# oracle/_ref is loaded only by the tests that use it (fixture `ref`), so a
# run that deselects them never maps it
needs_ref = pytest.mark.skipif(not os.path.exists(O.REF_LIB_PATH),
                               reason="oracle/_ref not built (reference tree absent)")


@pytest.fixture
def ref():
    r = O.ref_lib()
    if r is None:
        pytest.skip("oracle/_ref failed to load")
    return r


# ------------------------------------------------------------------ distance
@needs_ref
@pytest.mark.parametrize("d", [1, 3, 4, 5, 96, 128, 130, 960])
def test_dist_ref_matches_reference_header(d, ref):
    """wvo_dist_ref reproduces Distance::compare (distance.h:12-36) bit for bit."""
    g = np.random.default_rng(d)
    for _ in range(50):
        a = g.normal(size=d).astype(np.float32) * 3
        b = g.normal(size=d).astype(np.float32) * 3
        assert np.float32(O.dist_ref(a, b)).tobytes() == np.float32(ref.wvr_dist(a, b, d)).tobytes()


def test_dist_canon_hand_worked():
    # one chunk: ((1 + 4) + 9) + 16 = 30
    assert O.dist_canon(np.array([1, 2, 3, 4], np.float32), np.zeros(4, np.float32)) == 30.0
    # tail of 3: 1 + 4 + 9, no padding effect
    assert O.dist_canon(np.array([1, 2, 3], np.float32), np.zeros(3, np.float32)) == 14.0
    # 17 chunks: chunk 16 folds into accumulator 0; all-ones -> 4 per chunk
    assert O.dist_canon(np.ones(68, np.float32), np.zeros(68, np.float32)) == 68.0


def test_dist_canon_order_is_the_documented_tree():
    """A pattern whose float sum depends on association.  Chunk sums:
    s0 = 2^24, s1 = 1, s17 = 1.  Reference order: (2^24 + 1) + 1, each step
    rounding back to 2^24.  Build order: accumulator 1 = s1 + s17 = 2 exactly,
    the fold adds it to accumulator 0 in one step: 2^24 + 2 (representable)."""
    a = np.zeros(128, np.float32)
    a[0] = 4096.0
    a[4] = 1.0
    a[68] = 1.0
    b = np.zeros(128, np.float32)
    assert O.dist_ref(a, b) == 2.0 ** 24
    assert O.dist_canon(a, b) == 2.0 ** 24 + 2


@pytest.mark.parametrize("d", [96, 128, 960])
def test_dist_canon_vs_fp64(d):
    g = np.random.default_rng(d + 1)
    for _ in range(100):
        a = g.normal(size=d).astype(np.float32)
        b = g.normal(size=d).astype(np.float32)
        ref = O.dist_f64(a, b)
        assert abs(O.dist_canon(a, b) - ref) <= 1e-5 * ref
        assert abs(O.dist_ref(a, b) - O.dist_canon(a, b)) <= 1e-5 * ref


def test_dist_integer_data_exact_in_any_order():
    g = np.random.default_rng(5)
    for _ in range(100):
        a = g.integers(0, 256, 128).astype(np.float32)
        b = g.integers(0, 256, 128).astype(np.float32)
        exact = float(np.sum((a.astype(np.int64) - b.astype(np.int64)) ** 2))
        assert O.dist_ref(a, b) == exact
        assert O.dist_canon(a, b) == exact
        assert O.dist_f64(a, b) == exact


# ------------------------------------------------------------------ GenRandom
@needs_ref
@pytest.mark.parametrize("size,N", [(10, 10000), (100, 1000), (64, 65), (1, 2), (400, 2000000)])
def test_gen_random_algorithm_matches_reference(size, N, ref):
    """The oracle's GenRandom fed std::mt19937 draws equals util.h's GenRandom."""
    for seed in (1, 7, 12345):
        draws = np.zeros(size + 1, np.uint32)
        ref.wvr_mt19937(seed, draws, size + 1)
        want = np.zeros(size, np.uint32)
        ref.wvr_gen_random(seed, want, size, N)
        assert np.array_equal(O.gen_random_draws(draws, size, N), want)


@pytest.mark.parametrize("size,N", [(1, 2), (10, 11), (100, 10000), (400, 1000)])
def test_gen_random_invariants(size, N):
    for q in range(20):
        a = O.gen_random(O.stream(3, O.TAG_ENTRY, q), size, N)
        assert a.size == size and len(set(a.tolist())) == size and a.max() < N


def test_gen_random_rejects_size_ge_N():
    with pytest.raises(ValueError):
        O.gen_random(1, 5, 5)


# ------------------------------------------------------------------ exact kNN
def _numpy_knn(base, queries, k, self_offset=-1):
    b = base.astype(np.float64)
    out = []
    for i, q in enumerate(queries.astype(np.float64)):
        dd = ((b - q) ** 2).sum(1)
        order = np.lexsort((np.arange(len(b)), dd))
        if self_offset >= 0:
            order = order[order != i + self_offset]
        out.append(order[:k])
    return np.array(out, np.uint32)


def test_exact_knn_vs_numpy(small_float):
    base, q = small_float
    ids, dist = O.exact_knn(base[:500], q[:50], 10)
    assert np.array_equal(ids, _numpy_knn(base[:500], q[:50], 10))
    ids, _ = O.exact_knn(base[:300], base[:300], 7, self_offset=0)
    assert np.array_equal(ids, _numpy_knn(base[:300], base[:300], 7, self_offset=0))


def test_exact_knn_ties_by_id():
    base = np.zeros((6, 4), np.float32)   # all equidistant from the query
    ids, dist = O.exact_knn(base, np.ones((1, 4), np.float32), 4)
    assert ids.tolist() == [[0, 1, 2, 3]] and np.all(dist == 4.0)


# ------------------------------------------------------------------ metrics
def test_recall_hand_worked():
    gt = np.array([[0, 1, 2, 9], [3, 4, 5, 9]], np.uint32)
    res = np.array([[0, 1, 7], [5, 4, 3]], np.uint32)
    # misses: 7 in row 0 -> 1 of 6
    assert O.recall(res, gt, 3) == pytest.approx(1 - 1 / 6)


def test_graph_quality_hand_worked():
    ids = np.array([[1, 2, 9], [0, 9, 9]], np.uint32)
    deg = np.array([2, 1], np.uint32)
    gt = np.array([[1, 3], [2, 0]], np.uint32)
    # node 0: {1} of first 2 exact -> 1 / min(2, 2, 2); node 1: 0 of {2,0}? 0 is present -> 1 / min(2,1,2)
    assert O.graph_quality(ids, deg, gt, control=2) == pytest.approx((0.5 + 1.0) / 2)


# ------------------------------------------------------------------ greedy search KATs
def _chain_graph(n, order):
    g = np.full((n, 2), 0xFFFFFFFF, np.uint32)
    deg = np.zeros(n, np.uint32)
    for i in range(n):
        nb = [j for j in ((i + 1, i - 1) if order == "up" else (i - 1, i + 1)) if 0 <= j < n]
        g[i, :len(nb)] = nb
        deg[i] = len(nb)
    return g, deg


def kat_cases():
    """Hand-derived traces of RouteGreedy (component_route.cpp:16-47) on the
    line 0..7 with query at 10 (d_i = (10-i)^2)."""
    return [
        # graph i -> [i+1, i-1]; entries {0,1}; L=k=2: every expansion restarts at 0
        dict(order="up", entries=[0, 1], L=2, ids=[7, 6], dist=[9, 16], evals=10, hops=7),
        # graph i -> [i-1, i+1]; entries {4,5,0}; L=k=3: entries 4 and 5 are met
        # again as neighbours (not marked visited) and rejected as duplicates
        dict(order="down", entries=[4, 5, 0], L=3, ids=[7, 6, 5], dist=[9, 16, 25], evals=7, hops=3),
    ]


@pytest.mark.parametrize("case", kat_cases())
def test_search_kat(case):
    base = chain_points()
    g, deg = _chain_graph(8, case["order"])
    q = np.zeros((1, 4), np.float32)
    q[0, 0] = 10
    ent = np.array([case["entries"]], np.uint32)
    ids, dist, ev, hp = O.search_greedy(g, deg, base, q, case["L"], case["L"], entries=ent)
    assert ids[0].tolist() == case["ids"]
    assert dist[0].tolist() == case["dist"]
    assert ev[0] == case["evals"] and hp[0] == case["hops"]


def test_search_complete_graph_is_exact(small_float):
    base, q = small_float
    base = base[:200]
    n = len(base)
    g = np.array([[j for j in range(n) if j != i] for i in range(n)], np.uint32)
    deg = np.full(n, n - 1, np.uint32)
    gt, _ = O.exact_knn(base, q, 10)
    ids, _, _, _ = O.search_greedy(g, deg, base, q, 10, 10, seed=5)
    assert np.array_equal(ids, gt)


def test_search_all_entries_is_exact(small_float, small_float_graph):
    base, q = small_float
    ids_g, _, deg = small_float_graph
    n = 300
    sub = ids_g[:n].copy()
    sub[sub >= n] = 0xFFFFFFFF
    ent = np.tile(np.arange(n, dtype=np.uint32), (20, 1))
    gt, _ = O.exact_knn(base[:n], q[:20], 10)
    ids, _, _, _ = O.search_greedy(sub, np.minimum(deg[:n], 32), base[:n], q[:20], n, 10, entries=ent)
    assert np.array_equal(ids, gt)


def test_search_pool_invariants(small_float, small_float_graph):
    base, q = small_float
    ids_g, _, deg = small_float_graph
    ids, dist, ev, hp = O.search_greedy(ids_g, deg, base, q, 40, 40, seed=9)
    for r in range(len(q)):
        assert len(set(ids[r].tolist())) == 40
        keys = list(zip(dist[r].tolist(), ids[r].tolist()))
        assert keys == sorted(keys)
        for i in range(0, 40, 13):
            assert dist[r, i] == O.dist_canon(q[r], base[ids[r, i]])
    assert np.all(hp >= 1) and np.all(ev >= 40)


def test_search_rejects_bad_arguments(small_float, small_float_graph):
    base, q = small_float
    ids_g, _, deg = small_float_graph
    with pytest.raises(ValueError):   # L < k (src/builder.cpp:435-438)
        O.search_greedy(ids_g, deg, base, q, 5, 10)
    with pytest.raises(ValueError):   # L >= N with random entries
        O.search_greedy(ids_g, deg, base, q, len(base), 10)


# ------------------------------------------------------------------ init / NN-Descent
def test_init_random_invariants(small_int):
    base, _ = small_int
    ids, dist, cnt = O.init_random(base, 10, seed=3)
    for i in range(0, len(base), 97):
        row = ids[i, :cnt[i]]
        assert i not in row.tolist() and len(set(row.tolist())) == cnt[i] and cnt[i] in (9, 10)
        keys = list(zip(dist[i, :cnt[i]].tolist(), row.tolist()))
        assert keys == sorted(keys)
        assert all(dist[i, j] == O.dist_canon(base[i], base[row[j]]) for j in range(cnt[i]))


@pytest.mark.parametrize("slots", [0, 64])
def test_nndescent_full_init_is_exact_knn(small_float, slots):
    """S >= N-1 seeds every pool with all other nodes, so the result is the
    exact kNN graph (the ComponentInitKNNG closed form)."""
    base, _ = small_float
    base = base[:60]
    ids, dist, deg = O.nndescent(base, 8, 64, 59, 16, 3, seed=1, cand_slots=slots)
    gt, gd = O.exact_knn(base, base, 8, self_offset=0)
    assert np.array_equal(ids, gt) and np.all(deg == 8)


@pytest.mark.parametrize("slots", [0, 128])
def test_nndescent_invariants_and_quality(small_int, slots):
    base, _ = small_int
    ids, dist, deg = O.nndescent(base, 20, 30, 10, 30, 8, seed=4, cand_slots=slots)
    gt, _ = O.exact_knn(base, base, 10, self_offset=0)
    assert O.graph_quality(ids, deg, gt) > 0.95
    for i in range(0, len(base), 53):
        row = ids[i, :deg[i]].tolist()
        assert i not in row and len(set(row)) == len(row)
        keys = list(zip(dist[i, :deg[i]].tolist(), row))
        assert keys == sorted(keys)
        assert all(dist[i, j] == O.dist_canon(base[i], base[row[j]]) for j in range(deg[i]))


def test_nndescent_deterministic_across_threads(small_int):
    base, _ = small_int
    a = O.nndescent(base, 16, 24, 8, 24, 5, seed=2, cand_slots=64, nthreads=1)
    b = O.nndescent(base, 16, 24, 8, 24, 5, seed=2, cand_slots=64, nthreads=4)
    c = O.nndescent(base, 16, 24, 8, 24, 5, seed=2, cand_slots=0, nthreads=4)
    d = O.nndescent(base, 16, 24, 8, 24, 5, seed=2, cand_slots=0, nthreads=1)
    for x, y in ((a, b), (c, d)):
        assert all(np.array_equal(u, v) for u, v in zip(x, y))


def test_nndescent_iteration_zero_is_empty(small_int):
    """ITER counts (join, update) rounds including the empty first join."""
    base, _ = small_int
    _, _, _, st = O.nndescent(base, 16, 24, 8, 24, 3, seed=2, with_stats=True)
    assert st[0] == (0, 0) and st[1][0] > 0


def test_slot_mode_matches_exact_mode_quality(small_int):
    base, _ = small_int
    gt, _ = O.exact_knn(base, base, 10, self_offset=0)
    qe = O.graph_quality(*O.nndescent(base, 20, 30, 10, 30, 8, seed=6)[::2], gt)
    qs = O.graph_quality(*O.nndescent(base, 20, 30, 10, 30, 8, seed=6, cand_slots=128)[::2], gt)
    assert abs(qe - qs) < 0.005
