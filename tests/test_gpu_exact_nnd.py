"""Exact kNN (FP64) and NN-Descent on the GPU vs the CPU oracle."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


# ------------------------------------------------------------------ exact kNN
def test_exact_knn_bit_exact(ctx, small_float):
    base, q = small_float
    ids, dist, ms = ctx.exact_knn(base, q, 50)
    rid, rdist = O.exact_knn(base, q, 50)
    assert np.array_equal(ids, rid) and np.array_equal(dist, rdist)


def test_exact_knn_graph_self_excluded(ctx, small_int):
    base, _ = small_int
    ids, dist, _ = ctx.exact_knn(base, base[:300], 20, self_offset=0)
    rid, rdist = O.exact_knn(base, base[:300], 20, self_offset=0)
    assert np.array_equal(ids, rid) and np.array_equal(dist, rdist)
    assert not np.any(ids == np.arange(300)[:, None])


def test_exact_knn_d960(ctx):
    rng = np.random.default_rng(1)
    base = np.abs(rng.normal(size=(700, 960))).astype(np.float32)
    q = np.abs(rng.normal(size=(9, 960))).astype(np.float32)
    ids, dist, _ = ctx.exact_knn(base, q, 100)
    rid, rdist = O.exact_knn(base, q, 100)
    assert np.array_equal(ids, rid) and np.array_equal(dist, rdist)


# ------------------------------------------------------------------ NN-Descent
def _invariants(base, ids, dist, deg, K):
    N = len(base)
    assert np.all(deg <= K)
    for i in range(0, N, max(1, N // 97)):
        row = ids[i, :deg[i]].tolist()
        assert i not in row and len(set(row)) == len(row)
        keys = list(zip(dist[i, :deg[i]].tolist(), row))
        assert keys == sorted(keys)
        for j in range(0, deg[i], 7):
            assert dist[i, j] == O.dist_canon(base[i], base[row[j]])


@pytest.mark.parametrize("slots", [64, 256])
def test_nndescent_bit_exact_on_integer_data(ctx, small_int, slots):
    """On integer data every FP32 distance (MFMA Gram + norms included) is an
    exact integer, so the GPU build must equal the oracle's slot-table build
    (same seeds, same candidate-slot hash) bit for bit."""
    base, _ = small_int
    K, L, S, R, it = 24, 32, 10, 32, 6
    gi, gd, gdeg, st = ctx.nndescent_build(base, K, L, S, R, it, seed=3, cand_slots=slots)
    ri, rd, rdeg, rst = O.nndescent(base, K, L, S, R, it, seed=3, cand_slots=slots, with_stats=True)
    assert st["pairs"] == sum(p for p, _ in rst)
    assert st["proposals"] == sum(a for _, a in rst)
    assert np.array_equal(gdeg, rdeg)
    assert np.array_equal(gi, ri)
    assert np.array_equal(gd, rd)


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
def test_nndescent_grouped_order_bit_exact_on_integer_data(ctx, small_int, dtype):
    """With the nodes processed grouped by nearest pivot (forced on at this
    size) the build is still the oracle's, bit for bit: no per-node kernel's
    outcome depends on which wavefront serves a node, or when (BF16 of
    these small integers is exact too)."""
    base, _ = small_int
    K, L, S, R, it = 24, 32, 10, 32, 6
    gi, gd, gdeg, st = ctx.nndescent_build(base, K, L, S, R, it, seed=3, cand_slots=64, dtype=dtype, order=2)
    ri, rd, rdeg, rst = O.nndescent(base, K, L, S, R, it, seed=3, cand_slots=64, with_stats=True)
    assert st["pairs"] == sum(p for p, _ in rst)
    assert st["proposals"] == sum(a for _, a in rst)
    assert np.array_equal(gdeg, rdeg) and np.array_equal(gi, ri) and np.array_equal(gd, rd)


@pytest.mark.parametrize("dtype", ["f32", "bf16", "fp8"])
def test_nndescent_order_changes_speed_only(ctx, dtype):
    """60,000 continuous points (the Zipf mixture, where the automatic order
    is on): the graph built with nodes grouped by nearest pivot equals the
    one built in id order, in every join precision."""
    from weavess_amd import Index, datagen
    base, _ = datagen.gmm_c2(n_base=60_000, n_query=1)
    ix = Index(ctx, base)
    try:
        out = []
        for order in (1, 2):
            st = ix.build_nndescent(32, 48, 12, 64, 8, seed=5, dtype=dtype, cand_slots=128, order=order)
            out.append((ix.get_graph(), st["pairs"], st["proposals"]))
        (g1, p1, q1), (g2, p2, q2) = out
        assert p1 == p2 and q1 == q2
        for a_, b_ in zip(g1, g2):
            assert np.array_equal(a_, b_)
    finally:
        ix.close()


def test_nndescent_full_init_is_exact(ctx, small_float):
    base, _ = small_float
    base = base[:60]
    ids, dist, deg, _ = ctx.nndescent_build(base, 8, 64, 59, 16, 3, seed=1)
    gt, _ = O.exact_knn(base, base, 8, self_offset=0)
    assert np.array_equal(ids, gt)


@pytest.mark.parametrize("dtype", ["f32", "bf16", "fp8"])
def test_nndescent_quality_and_invariants(ctx, small_float, dtype):
    base, _ = small_float
    K, L, S, R, it = 20, 32, 10, 32, 8
    ids, dist, deg, st = ctx.nndescent_build(base, K, L, S, R, it, seed=5, dtype=dtype)
    _invariants(base, ids, dist, deg, K)
    gt, _ = O.exact_knn(base, base, 10, self_offset=0)
    q_gpu = O.graph_quality(ids, deg, gt)
    ri, rd, rdeg = O.nndescent(base, K, L, S, R, it, seed=5)
    q_ref = O.graph_quality(ri, rdeg, gt)
    print(dtype, "graph recall gpu", q_gpu, "oracle", q_ref, st)
    tol = 0.005 if dtype == "f32" else 0.02
    assert q_gpu >= q_ref - tol


def test_nndescent_deterministic(ctx, small_float):
    base, _ = small_float
    a = ctx.nndescent_build(base, 16, 24, 8, 24, 5, seed=2)
    b = ctx.nndescent_build(base, 16, 24, 8, 24, 5, seed=2)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_nndescent_rejects_bad_params(ctx, small_float):
    from weavess_amd import WvError
    base, _ = small_float
    for bad in [dict(K=40, L=32), dict(S=0), dict(S=len(base))]:
        p = dict(K=16, L=32, S=8, R=16)
        p.update(bad)
        with pytest.raises(WvError):
            ctx.nndescent_build(base, p["K"], p["L"], p["S"], p["R"], 3)


def test_monitor_per_round_matches_oracle(ctx, small_int):
    """The per-round graph recall of the debug monitor (component_refine.cpp:
    106-118, 241-275) against the oracle's pools after each round, on
    integer data where the builds are bit-exact; the monitored build's graph
    equals the plain build's."""
    from weavess_amd import Index
    base, _ = small_int
    K, L, S, R, it, slots = 24, 32, 10, 32, 6, 64
    ctrl = np.random.default_rng(4).choice(len(base), 100, replace=False).astype(np.uint32)
    truth, _ = O.exact_knn(base, base[ctrl], 100)  # top 100 including the node itself, as the reference
    ix = Index(ctx, base)
    st, per = ix.build_nndescent_monitor(K, L, S, R, it, ctrl, truth, seed=3, cand_slots=slots)
    g, _, gdeg = ix.get_graph()
    ix.close()
    ids, _, deg, _ = ctx.nndescent_build(base, K, L, S, R, it, seed=3, cand_slots=slots)
    assert np.array_equal(g, ids) and np.array_equal(gdeg, deg)
    want = []
    for r in range(1, it + 1):
        s = O.nndescent_state(base, K, L, S, R, r, seed=3, cand_slots=slots, nthreads=0)
        hits = [len(set(s["pool_ids"][c, :s["pool_cnt"][c]].tolist()) & set(truth[i].tolist()))
                for i, c in enumerate(ctrl)]
        want.append(sum(h / 100 for h in hits) / len(ctrl))
    np.testing.assert_allclose(per, want, rtol=0, atol=1e-12)
    assert per[-1] > per[0]
