"""Edge cases on the GPU against the oracle: empty and ragged inputs, the
largest shapes the entry points accept, degenerate data (equal rows, so
every distance ties and ids decide the order)."""
import numpy as np
import pytest

import oracle as O
from weavess_amd import Index

pytestmark = pytest.mark.gpu


def _same(got, ref, counts=True):
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    if counts:
        assert np.array_equal(got[2], ref[2]) and np.array_equal(got[3], ref[3])


def _int_data(n, d, nq, seed):
    g = np.random.default_rng(seed)
    return g.integers(0, 16, (n, d)).astype(np.float32), g.integers(0, 16, (nq, d)).astype(np.float32)


def test_empty_query_batch(ctx, small_float, small_float_graph):
    base, q = small_float
    g, _, deg = small_float_graph
    ids, dist, ev, hp, st = ctx.search_greedy(g, deg, base, q[:0], 32, 10, seed=1)
    assert ids.shape == (0, 10) and st["hops"] == 0


def test_ragged_graph(ctx):
    """Isolated nodes (degree 0), padding slots holding ids >= N, degrees
    below the row width: skipped exactly as the oracle skips them."""
    base, q = _int_data(500, 16, 40, 3)
    g = np.random.default_rng(4)
    deg = g.integers(0, 9, 500).astype(np.uint32)
    deg[:50] = 0
    graph = np.full((500, 8), 0xFFFFFFFF, np.uint32)
    for i in range(500):
        nb = g.choice(500, 8, replace=False).astype(np.uint32)
        nb[g.uniform(size=8) < 0.15] = 500 + g.integers(0, 100)  # out-of-range ids are skipped
        graph[i] = nb
    ref = O.search_greedy(graph, deg, base, q, 24, 10, seed=2)
    got = ctx.search_greedy(graph, deg, base, q, 24, 10, seed=2)
    _same(got, ref)


@pytest.mark.parametrize("d", [1, 1024])
def test_smallest_and_largest_dimension(ctx, d):
    base, q = _int_data(600, d, 30, 5)
    gi, _, gdeg = O.nndescent(base, 12, 20, 8, 20, 4, seed=2)
    ref = O.search_greedy(gi, gdeg, base, q, 40, 10, seed=6)
    got = ctx.search_greedy(gi, gdeg, base, q, 40, 10, seed=6)
    _same(got, ref, counts=False)
    assert np.array_equal(got[3], ref[3]) and np.all(got[2] >= ref[2])  # the default table may forget
    ix = Index(ctx, base)
    ix.set_graph(gi, gdeg)
    ix.set_search_config(visited_slots=16384)  # holds every visited id: exact counts
    got = ix.search(q, 40, 10, seed=6)
    ix.close()
    _same(got, ref)


def test_largest_random_pool_and_whole_base_entries(ctx):
    """L = N-1 random entries (GenRandom's limit), and all N nodes given as
    entries: the pool then holds the exact kNN."""
    base, q = _int_data(300, 8, 20, 7)
    gi, _, gdeg = O.nndescent(base, 10, 16, 6, 16, 4, seed=3)
    ref = O.search_greedy(gi, gdeg, base, q, 299, 20, seed=9)
    got = ctx.search_greedy(gi, gdeg, base, q, 299, 20, seed=9)
    _same(got, ref)
    ent = np.tile(np.arange(300, dtype=np.uint32), (len(q), 1))
    got = ctx.search_greedy(gi, gdeg, base, q, 300, 300, entries=ent)
    ex, exd = O.exact_knn(base, q, 128)
    assert np.array_equal(got[0][:, :128], ex)


def test_largest_pool(ctx):
    """L = k = 4096, the largest pool, against the oracle on 5,000 points."""
    base, q = _int_data(5000, 8, 6, 11)
    gi, _, gdeg = O.nndescent(base, 16, 24, 8, 24, 4, seed=5)
    ref = O.search_greedy(gi, gdeg, base, q, 4096, 4096, seed=3)
    ix = Index(ctx, base)
    ix.set_graph(gi, gdeg)
    ix.set_search_config(visited_slots=16384)  # the largest table that fits beside this pool
    got = ix.search(q, 4096, 4096, seed=3)
    ix.close()
    # ~4,900 ids in 16,384 slots: a probe run past 8 slots can still forget,
    # which re-evaluates (evaluations only grow); everything else is exact
    _same(got, ref, counts=False)
    assert np.array_equal(got[3], ref[3]) and np.all(got[2] >= ref[2])
    from weavess_amd import WvError
    with pytest.raises(WvError):
        ctx.search_greedy(gi, gdeg, base, q, 4097, 10, seed=3)


def test_equal_rows_tie_by_id(ctx):
    """Every base row equal: all distances tie, and (distance, id) order makes
    the result the smallest ids reached, bit-exact with the oracle."""
    base = np.ones((400, 12), np.float32)
    q = np.zeros((10, 12), np.float32)
    gi, _, gdeg = O.nndescent(base, 8, 12, 6, 12, 3, seed=1)
    ref = O.search_greedy(gi, gdeg, base, q, 30, 10, seed=4)
    got = ctx.search_greedy(gi, gdeg, base, q, 30, 10, seed=4)
    _same(got, ref)
    assert np.all(got[1] == 12.0)


def test_nndescent_tiny_and_degenerate(ctx):
    """N = 24 with S = 10 (several rounds over almost all pairs) and a base of
    equal rows: the GPU build equals the oracle's slot-table build bit for
    bit on integer data."""
    base, _ = _int_data(24, 8, 1, 13)
    for b in (base, np.ones((40, 8), np.float32)):
        gi, gd, gdeg, st = ctx.nndescent_build(b, 6, 10, 10, 10, 4, seed=2, cand_slots=64)
        ri, rd, rdeg = O.nndescent(b, 6, 10, 10, 10, 4, seed=2, cand_slots=64)
        assert np.array_equal(gdeg, rdeg) and np.array_equal(gi, ri) and np.array_equal(gd, rd)


def test_exact_knn_largest_k(ctx):
    base, q = _int_data(1000, 16, 25, 17)
    got, gd, _ = ctx.exact_knn(base, q, 128)
    ref, rd = O.exact_knn(base, q, 128)
    assert np.array_equal(got, ref) and np.array_equal(gd, rd)
