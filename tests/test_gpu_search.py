"""Greedy search on the GPU vs the CPU oracle (parity: bit-exact ids).

The oracle and the kernel share the (dist, id) order and the build-order
FP32 distance, so ids AND distances must match exactly; evaluation and hop
counts must match whenever the visited table does not overflow.
"""
import numpy as np
import pytest

import oracle as O
from tests.conftest import chain_points
from tests.test_oracle import _chain_graph, kat_cases

pytestmark = pytest.mark.gpu


def _gpu_search(ctx, graph, deg, base, q, L, k, **kw):
    ids, dist, ev, hp, st = ctx.search_greedy(graph, deg, base, q, L, k, **kw)
    return ids, dist, ev, hp


@pytest.mark.parametrize("L", [10, 20, 64, 100, 200])
def test_search_parity_integer(ctx, small_int, small_int_graph, L):
    base, q = small_int
    g, _, deg = small_int_graph
    ref = O.search_greedy(g, deg, base, q, L, 10, seed=21)
    got = _gpu_search(ctx, g, deg, base, q, L, 10, seed=21)
    print("L", L, "evals", got[2].mean())
    assert np.array_equal(got[0], ref[0])
    assert np.array_equal(got[1], ref[1])
    assert np.array_equal(got[2], ref[2]) and np.array_equal(got[3], ref[3])


@pytest.mark.parametrize("L", [16, 48, 130])
def test_search_parity_float(ctx, small_float, small_float_graph, L):
    base, q = small_float
    g, _, deg = small_float_graph
    ref = O.search_greedy(g, deg, base, q, L, L, seed=5)
    got = _gpu_search(ctx, g, deg, base, q, L, L, seed=5)
    assert np.array_equal(got[0], ref[0])
    # same build-order arithmetic: distances agree bit for bit (<= 1e-5 rel is the stated bar)
    np.testing.assert_allclose(got[1], ref[1], rtol=1e-5, atol=0)
    assert np.array_equal(got[1], ref[1])
    assert np.array_equal(got[2], ref[2]) and np.array_equal(got[3], ref[3])


@pytest.mark.parametrize("case", kat_cases())
def test_search_kat_gpu(ctx, case):
    base = chain_points()
    g, deg = _chain_graph(8, case["order"])
    q = np.zeros((1, 4), np.float32)
    q[0, 0] = 10
    ent = np.array([case["entries"]], np.uint32)
    ids, dist, ev, hp = _gpu_search(ctx, g, deg, base, q, case["L"], case["L"], entries=ent)
    assert ids[0].tolist() == case["ids"] and dist[0].tolist() == case["dist"]
    assert ev[0] == case["evals"] and hp[0] == case["hops"]


def test_search_explicit_entries_and_strides(ctx, small_float, small_float_graph):
    base, q = small_float
    g, _, deg = small_float_graph
    rng = np.random.default_rng(3)
    ent = np.stack([rng.choice(len(base), 40, replace=False) for _ in range(len(q))]).astype(np.uint32)
    ref = O.search_greedy(g, deg, base, q, 40, 10, entries=ent, k_stride=12)
    got = _gpu_search(ctx, g, deg, base, q, 40, 10, entries=ent, k_stride=12)
    assert np.array_equal(got[0], ref[0]) and np.all(got[0][:, 10:] == 0xFFFFFFFF)


def test_search_per_query_budgets(ctx, small_float, small_float_graph):
    base, q = small_float
    g, _, deg = small_float_graph
    rng = np.random.default_rng(4)
    Lq = rng.integers(16, 300, len(q)).astype(np.uint32)
    kq = rng.choice([1, 10, 100], len(q)).astype(np.uint32)
    Lq = np.maximum(Lq, kq)
    ref = O.search_greedy(g, deg, base, q, 0, 0, seed=8, L_per_q=Lq, k_per_q=kq)
    got = _gpu_search(ctx, g, deg, base, q, 0, 0, seed=8, L_per_q=Lq, k_per_q=kq)
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])


@pytest.mark.parametrize("d", [96, 99, 960])
def test_search_other_dims(ctx, d):
    rng = np.random.default_rng(d)
    base = rng.normal(size=(800, d)).astype(np.float32)
    q = rng.normal(size=(40, d)).astype(np.float32)
    g, _, deg = O.nndescent(base, 16, 24, 8, 24, 5, seed=1)
    ref = O.search_greedy(g, deg, base, q, 32, 10, seed=2)
    got = _gpu_search(ctx, g, deg, base, q, 32, 10, seed=2)
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])


@pytest.mark.parametrize("d", [68, 72, 96])
@pytest.mark.parametrize("rg", [4, 8, 16])
def test_search_eight_lane_rows(ctx, d, rg):
    """Rows of 17-24 chunks (d = 65..96) are gathered 8 lanes a row, 3 chunks
    a lane, each lane's two accumulators summed as the fold's first step:
    ids, distances, evaluations and hops equal the oracle's for every gather
    tile, on the grid and persistent schedules, with 2 wavefronts per query
    and with the merge buffer.  d = 68 and 72 leave the last chunk slots of
    some lanes empty."""
    from weavess_amd import Index
    rng = np.random.default_rng(d + rg)
    base = rng.normal(size=(1500, d)).astype(np.float32)
    q = rng.normal(size=(60, d)).astype(np.float32)
    g, _, deg = O.nndescent(base, 24, 32, 8, 32, 5, seed=1)
    ref = O.search_greedy(g, deg, base, q, 200, 10, seed=4)
    ix = Index(ctx, base)
    ix.set_graph(g, deg)
    for extra in ({}, {"schedule": 2}, {"waves_per_query": 2}, {"merge_buffer": 64}):
        # a table ten times the base never forgets, so the counts are exact
        ix.set_search_config(row_groups=rg, visited_slots=16384, **extra)
        got = ix.search(q, 200, 10, seed=4)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), extra
        assert np.array_equal(got[2], ref[2]), ("evals", extra)
        assert np.array_equal(got[3], ref[3]), ("hops", extra)
    ix.close()


def test_search_complete_graph_exact(ctx, small_float):
    base, q = small_float
    base = base[:200]
    n = len(base)
    g = np.array([[j for j in range(n) if j != i] for i in range(n)], np.uint32)
    deg = np.full(n, n - 1, np.uint32)
    gt, _ = O.exact_knn(base, q, 10)
    ids, *_ = _gpu_search(ctx, g, deg, base, q, 10, 10, seed=5)
    assert np.array_equal(ids, gt)


def test_search_config_changes_speed_only(ctx, small_float, small_float_graph):
    """waves per block, gather tile and a tiny (forgetful) visited table
    change nothing but the evaluation count."""
    from weavess_amd import Index
    base, q = small_float
    g, _, deg = small_float_graph
    ix = Index(ctx, base)
    ix.set_graph(g, deg)
    ref = ix.search(q, 64, 10, seed=3)
    for cfg in [dict(waves_per_block=2), dict(waves_per_block=4, row_groups=4),
                dict(row_groups=16), dict(visited_slots=64), dict(merge_buffer=64),
                dict(merge_buffer=128, visited_slots=64, waves_per_block=2)]:
        ix.set_search_config(**cfg)
        got = ix.search(q, 64, 10, seed=3)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), cfg
        if cfg.get("visited_slots"):
            assert np.all(got[2] >= ref[2])      # a forgetting table only re-evaluates
        else:
            assert np.array_equal(got[2], ref[2]), cfg
        assert np.array_equal(got[3], ref[3])
    ix.close()


@pytest.mark.parametrize("L", [40, 700])
def test_search_autotune_changes_speed_only(ctx, small_float, small_float_graph, L):
    """The autotuner keeps a valid tuning in the index; ids, distances and
    hops are unchanged (evaluations only grow where a smaller visited table
    forgets)."""
    from weavess_amd import Index
    base, q = small_float
    g, _, deg = small_float_graph
    ref = O.search_greedy(g, deg, base, q, L, 10, seed=6)
    ix = Index(ctx, base)
    ix.set_graph(g, deg)
    cfg, ms = ix.autotune(q, L, 10, seed=6)
    assert ms > 0 and cfg["row_groups"] in (0, 4, 8, 16) and cfg["waves_per_block"] in (0, 1, 2, 4)
    assert cfg["visited_slots"] in (0, 256, 512, 1024, 2048, 4096)
    got = ix.search(q, L, 10, seed=6)
    ix.close()
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    assert np.array_equal(got[3], ref[3]) and np.all(got[2] >= ref[2])


def test_search_autotune_budgets_changes_speed_only(ctx, small_float, small_float_graph):
    """The autotuner under per-query budgets keeps a valid tuning in the
    index, and the uneven-budget search it serves still returns the oracle's
    ids, distances and hops (evaluations only grow where a smaller visited
    table forgets)."""
    from weavess_amd import Index
    from weavess_amd.devmem import DeviceArray, pad_rows
    base, q = small_float
    g, _, deg = small_float_graph
    rng = np.random.default_rng(11)
    Lq = rng.integers(16, 300, size=len(q)).astype(np.uint32)
    kq = rng.choice(np.array([1, 10, 16], np.uint32), size=len(q))
    ref = O.search_greedy(g, deg, base, q, 0, 0, seed=6, L_per_q=Lq, k_per_q=kq, k_stride=16)
    ix = Index(ctx, base)
    ix.set_graph(g, deg)
    dq, dL, dk = DeviceArray.from_host(pad_rows(q)), DeviceArray.from_host(Lq), DeviceArray.from_host(kq)
    cfg, ms = ix.autotune_budgets_dev(dq.ptr, len(q), dL.ptr, dk.ptr, 16, seed=6)
    for a_ in (dq, dL, dk):
        a_.free()
    assert ms > 0 and cfg["schedule"] in (0, 1, 2) and cfg["waves_per_query"] in (0, 1, 2, 4)
    got = ix.search(q, 0, 0, seed=6, L_per_q=Lq, k_per_q=kq, k_stride=16)
    ix.close()
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    assert np.array_equal(got[3], ref[3]) and np.all(got[2] >= ref[2])


@pytest.mark.parametrize("L", [48, 130, 600, 1500])
@pytest.mark.parametrize("mb", [1, 64, 256])
def test_search_merge_buffer_parity(ctx, small_float, small_float_graph, L, mb):
    """The side buffer that delays pool merges (none, 64 or 256 keys) leaves
    ids, distances, evaluations and hops bit-identical to the oracle, at small
    pools and at pools most of the base (a visited table large enough for the
    whole base keeps the evaluation counts exact)."""
    from weavess_amd import Index
    base, q = small_float
    g, _, deg = small_float_graph
    ref = O.search_greedy(g, deg, base, q, L, 10, seed=8)
    ix = Index(ctx, base)
    ix.set_graph(g, deg)
    ix.set_search_config(merge_buffer=mb, visited_slots=16384)
    got = ix.search(q, L, 10, seed=8)
    ix.close()
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    assert np.array_equal(got[2], ref[2]) and np.array_equal(got[3], ref[3])


@pytest.mark.parametrize("L", [100, 700, 1900])
def test_search_merge_buffer_integer_ties(ctx, small_int, small_int_graph, L):
    """Integer data (many equal distances, ordered by id) through the default
    plan, which buffers merges from L = 384 up; k = L returns the whole pool."""
    from weavess_amd import Index
    base, q = small_int
    g, _, deg = small_int_graph
    ref = O.search_greedy(g, deg, base, q[:50], L, L, seed=21)
    ix = Index(ctx, base)
    ix.set_graph(g, deg)
    got = ix.search(q[:50], L, L, seed=21)
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    assert np.array_equal(got[3], ref[3]) and np.all(got[2] >= ref[2])  # default table may forget
    ix.set_search_config(visited_slots=16384)
    got = ix.search(q[:50], L, L, seed=21)
    ix.close()
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    assert np.array_equal(got[2], ref[2]) and np.array_equal(got[3], ref[3])


def _many_queries(base, n, seed):
    rng = np.random.default_rng(seed)
    return (base[rng.integers(0, len(base), n)] + rng.normal(0, 0.05, (n, base.shape[1]))).astype(np.float32)


@pytest.mark.parametrize("wpb", [1, 2, 4])
def test_search_persistent_queue_many_queries(ctx, small_float, small_float_graph, wpb):
    """The persistent schedule (wavefronts taking queries from the device work
    queue): 40,000 queries, several per wavefront.  Every query is served
    once, with the grid schedule's ids, distances, evaluations and hops, and
    the oracle's ids, distances and hops."""
    from weavess_amd import Index
    base, _ = small_float
    g, _, deg = small_float_graph
    qq = _many_queries(base, 40_000, 9)
    ix = Index(ctx, base)
    ix.set_graph(g, deg)
    ix.set_search_config(waves_per_block=wpb, schedule=1)
    grid = ix.search(qq, 96, 10, seed=4)
    ix.set_search_config(waves_per_block=wpb, schedule=2)
    got = ix.search(qq, 96, 10, seed=4)
    ix.close()
    for i in range(4):
        assert np.array_equal(got[i], grid[i]), i
    ref = O.search_greedy(g, deg, base, qq[:3000], 96, 10, seed=4)
    assert np.array_equal(got[0][:3000], ref[0]) and np.array_equal(got[1][:3000], ref[1])
    assert np.array_equal(got[3][:3000], ref[3])


def test_search_persistent_budgets_and_model(ctx, small_float, small_float_graph):
    """Per-query L and k (slots served largest L first) and the learned stop
    under the persistent schedule equal the grid schedule and the oracle."""
    from weavess_amd import Index
    from weavess_amd._capi import TermModel
    base, _ = small_float
    g, _, deg = small_float_graph
    qq = _many_queries(base, 20_000, 3)
    rng = np.random.default_rng(5)
    Lq = rng.integers(16, 300, len(qq)).astype(np.uint32)
    kq = np.minimum(rng.choice([1, 10, 100], len(qq)), Lq).astype(np.uint32)
    m = TermModel.init(3)
    m.b2 = -1.0
    m.logit_threshold = 0.0
    ix = Index(ctx, base)
    ix.set_graph(g, deg)
    out = {}
    for sched in (1, 2):
        ix.set_search_config(schedule=sched)
        out[sched] = (ix.search(qq, 0, 0, seed=6, L_per_q=Lq, k_per_q=kq, k_stride=100),
                      ix.search_ext(qq, 150, 10, seed=6, model=m))
    ix.close()
    for i in range(4):
        assert np.array_equal(out[2][0][i], out[1][0][i]), i
        assert np.array_equal(out[2][1][i], out[1][1][i]), i
    n = 2000
    ref = O.search_greedy(g, deg, base, qq[:n], 0, 0, seed=6, L_per_q=Lq[:n], k_per_q=kq[:n], k_stride=100)
    assert np.array_equal(out[2][0][0][:n], ref[0]) and np.array_equal(out[2][0][3][:n], ref[3])


def test_search_persistent_graph_replays_and_noalloc(ctx, small_float, small_float_graph):
    """The persistent schedule inside a captured launch (the counter is zeroed
    by a kernel of the launch, so every replay serves every query) and in the
    no-allocation entry point (its counter in the library's array)."""
    from weavess_amd import Index
    from weavess_amd.devmem import DeviceArray, pad_rows
    base, _ = small_float
    g, _, deg = small_float_graph
    qq = _many_queries(base, 30_000, 11)
    Q = len(qq)
    ix = Index(ctx, base)
    ix.set_graph(g, deg)
    ix.set_search_config(schedule=1)
    ref = ix.search(qq, 64, 10, seed=2)
    ix.set_search_config(schedule=2)
    dq = DeviceArray.from_host(pad_rows(qq))
    d_ids, d_ev = DeviceArray((Q, 10), np.uint32), DeviceArray((Q,), np.uint32)
    sg = ix.search_graph(dq.ptr, Q, 64, 10, d_ids.ptr, seed=2, d_evals=d_ev.ptr)
    for rep in range(3):
        ix.search(qq[:100], 32, 5, seed=1)  # eager work between replays
        sg.launch()
        assert np.array_equal(d_ids.to_host(), ref[0]) and np.array_equal(d_ev.to_host(), ref[2]), rep
    sg.close()
    for n in (Q, 5000):
        st = ix.search_dev_noalloc(dq.ptr, n, 64, 10, d_ids.ptr, seed=2, d_evals=d_ev.ptr)
        assert np.array_equal(d_ids.to_host()[:n], ref[0][:n]) and np.array_equal(d_ev.to_host()[:n], ref[2][:n])
        assert st["dist_evals"] == int(ref[2][:n].sum()) and st["hops"] == int(ref[3][:n].sum())
    for a_ in (dq, d_ids, d_ev):
        a_.free()
    ix.close()


@pytest.mark.parametrize("wpq", [2, 4])
@pytest.mark.parametrize("L,mb", [(48, 1), (300, 128), (1500, 256)])
def test_search_waves_per_query_parity(ctx, small_float, small_float_graph, wpq, L, mb):
    """Several wavefronts per query (the neighbours' visited tests and
    distances split between them, wave 0 merging block by block) return the
    oracle's ids, distances, evaluations and hops (a table large enough to
    keep the counts exact), with random and with explicit entries."""
    from weavess_amd import Index
    base, q = small_float
    g, _, deg = small_float_graph
    ref = O.search_greedy(g, deg, base, q, L, 10, seed=8)
    ix = Index(ctx, base)
    ix.set_graph(g, deg)
    ix.set_search_config(merge_buffer=mb, visited_slots=8192, waves_per_query=wpq)
    got = ix.search(q, L, 10, seed=8)
    for i in range(4):
        assert np.array_equal(got[i], ref[i]), i
    ent = np.stack([np.random.default_rng(r).choice(len(base), L, replace=False) for r in range(len(q))])
    ent = ent.astype(np.uint32)
    ref2 = O.search_greedy(g, deg, base, q, L, L, entries=ent)
    got2 = ix.search(q, L, L, entries=ent)
    ix.close()
    for i in range(4):
        assert np.array_equal(got2[i], ref2[i]), i


@pytest.mark.parametrize("d,tiles", [(16, (8, 16)), (128, (4, 8, 16)), (96, (4, 8, 16)), (200, (4, 8)),
                                     (512, (8,)), (960, (8,))])
def test_search_look_ahead_loop(ctx, d, tiles):
    """Launches whose LDS allows at most 2 wavefronts a SIMD (a 16,384-slot
    table, or a pool of about 2,000 keys) on graphs of at most 64 slots a
    node take the loop that tests the next node's neighbours and issues
    their rows before the current node's merge: ids, distances, evaluations
    and hops equal the oracle's for every gather tile, with and without the
    merge buffer; a degree-80 graph under the same LDS keeps the loop without
    look-ahead and equals the oracle too."""
    from weavess_amd import Index
    rng = np.random.default_rng(d)
    base = rng.normal(size=(2500, d)).astype(np.float32)
    q = rng.normal(size=(40, d)).astype(np.float32)
    g, _, deg = O.nndescent(base, 24, 32, 8, 32, 4, seed=1)
    ix = Index(ctx, base)
    ix.set_graph(g, deg)
    for L, vs in ((200, 16384), (1990, 0)):
        ref = O.search_greedy(g, deg, base, q, L, 10, seed=6)
        for rg in tiles:
            for mb in (1, 64):
                ix.set_search_config(row_groups=rg, visited_slots=vs, merge_buffer=mb)
                got = ix.search(q, L, 10, seed=6)
                for i in (0, 1, 3):
                    assert np.array_equal(got[i], ref[i]), (L, rg, mb, i)
                if vs:  # a table larger than the base forgets nothing
                    assert np.array_equal(got[2], ref[2]), (L, rg, mb)
                else:
                    assert np.all(got[2] >= ref[2])
    ix.close()
    g80, _ = O.exact_knn(base, base, 80, self_offset=0)
    deg80 = np.full(len(base), 80, np.uint32)
    ref = O.search_greedy(g80, deg80, base, q[:10], 200, 10, seed=6)
    ix = Index(ctx, base)
    ix.set_graph(g80, deg80)
    ix.set_search_config(visited_slots=16384)
    got = ix.search(q[:10], 200, 10, seed=6)
    ix.close()
    for i in range(4):
        assert np.array_equal(got[i], ref[i]), i


@pytest.mark.parametrize("wpq", [2, 4])
def test_search_waves_per_query_wide_rows_and_degrees(ctx, wpq):
    """d = 960 rows (16 chunks a lane) and a degree-128 graph (more than one
    wavefront's worth of neighbours a hop), per-query L and k."""
    from weavess_amd import Index
    rng = np.random.default_rng(3)
    base = np.abs(rng.normal(size=(3000, 960))).astype(np.float32)
    q = np.abs(rng.normal(size=(120, 960))).astype(np.float32)
    g, _ = O.exact_knn(base, base, 128, self_offset=0)
    deg = np.full(len(base), 128, np.uint32)
    Lq = rng.integers(16, 400, len(q)).astype(np.uint32)
    kq = np.minimum(rng.choice([1, 10, 64], len(q)), Lq).astype(np.uint32)
    ref = O.search_greedy(g, deg, base, q, 0, 0, seed=2, L_per_q=Lq, k_per_q=kq, k_stride=64)
    ix = Index(ctx, base)
    ix.set_graph(g, deg)
    ix.set_search_config(visited_slots=16384, waves_per_query=wpq)
    got = ix.search(q, 0, 0, seed=2, L_per_q=Lq, k_per_q=kq, k_stride=64)
    ix.close()
    for i in range(4):
        assert np.array_equal(got[i], ref[i]), i


def test_search_rejects_bad_arguments(ctx, small_float, small_float_graph):
    from weavess_amd import WvError
    base, q = small_float
    g, _, deg = small_float_graph
    with pytest.raises(WvError) as e:
        _gpu_search(ctx, g, deg, base, q, 5, 10)  # L < k
    assert e.value.code == -22
    with pytest.raises(WvError):
        _gpu_search(ctx, g, deg, base, q, len(base), 10)  # L >= N, random entries


def test_search_graph_replays_match_eager(ctx, small_float, small_float_graph):
    """The captured launch, replayed with the program's own eager searches and
    synchronisations in between and the query rows rewritten between replays,
    returns on every replay exactly what the eager search returns."""
    from weavess_amd import Index
    from weavess_amd.devmem import DeviceArray, pad_rows
    base, q = small_float
    g, _, deg = small_float_graph
    ix = Index(ctx, base)
    ix.set_graph(g, deg)
    q1, q2 = q, q[::-1].copy()
    refs = [ix.search(q1, 64, 10, seed=4), ix.search(q2, 64, 10, seed=4)]
    Q = len(q)
    dq = DeviceArray.from_host(pad_rows(q1))
    d_ids, d_dist = DeviceArray((Q, 10), np.uint32), DeviceArray((Q, 10), np.float32)
    d_ev, d_hp = DeviceArray((Q,), np.uint32), DeviceArray((Q,), np.uint32)
    sg = ix.search_graph(dq.ptr, Q, 64, 10, d_ids.ptr, seed=4, d_out_dist=d_dist.ptr, d_evals=d_ev.ptr,
                         d_hops=d_hp.ptr)
    for rep in range(4):
        ix.search(q2[:50], 32, 5, seed=1)  # eager work between replays
        dq.copy_from_host(pad_rows(q1 if rep % 2 == 0 else q2))
        sg.launch()
        ref = refs[rep % 2]
        assert np.array_equal(d_ids.to_host(), ref[0]) and np.array_equal(d_dist.to_host(), ref[1]), rep
        assert np.array_equal(d_ev.to_host(), ref[2]) and np.array_equal(d_hp.to_host(), ref[3]), rep
    sg.close()
    for a in (dq, d_ids, d_dist, d_ev, d_hp):
        a.free()
    ix.close()


def test_search_graph_per_query_budgets_and_model(ctx, small_float, small_float_graph):
    """Per-query L and k, and the learned stop, through the captured launch."""
    from weavess_amd import Index
    from weavess_amd._capi import TermModel
    from weavess_amd.devmem import DeviceArray, pad_rows
    base, q = small_float
    g, _, deg = small_float_graph
    ix = Index(ctx, base)
    ix.set_graph(g, deg)
    Q = len(q)
    rng = np.random.default_rng(2)
    Lq = rng.integers(16, 200, Q).astype(np.uint32)
    kq = np.minimum(rng.choice([1, 10, 16], Q), Lq).astype(np.uint32)
    ref = ix.search(q, 0, 0, seed=5, L_per_q=Lq, k_per_q=kq, k_stride=16)
    m = TermModel.init(3)
    m.b2 = -1.0
    m.logit_threshold = 0.0
    mref = ix.search_ext(q, 120, 10, seed=5, model=m)
    dq = DeviceArray.from_host(pad_rows(q))
    dL, dk = DeviceArray.from_host(Lq), DeviceArray.from_host(kq)
    d_ids, d_ev = DeviceArray((Q, 16), np.uint32), DeviceArray((Q,), np.uint32)
    sg = ix.search_graph(dq.ptr, Q, 0, 0, d_ids.ptr, seed=5, d_L=dL.ptr, d_k=dk.ptr, k_stride=16, d_evals=d_ev.ptr)
    m_ids = DeviceArray((Q, 10), np.uint32)
    sm = ix.search_graph(dq.ptr, Q, 120, 10, m_ids.ptr, seed=5, model=m)
    for _ in range(3):
        sg.launch()
        sm.launch()
        assert np.array_equal(d_ids.to_host(), ref[0]) and np.array_equal(d_ev.to_host(), ref[2])
        assert np.array_equal(m_ids.to_host(), mref[0])
    sg.close()
    sm.close()
    for a in (dq, dL, dk, d_ids, d_ev, m_ids):
        a.free()
    ix.close()


def test_gather_rate(ctx):
    """The gather-rate measurement runs the search's gather tile over random
    rows at a fixed residency and reports a positive, finite rate; more
    resident wavefronts do not make it slower by more than noise; shapes
    outside its range are refused."""
    from weavess_amd import WvError
    lo, ms_lo = ctx.gather_rate(200_000, 96, 2, batches=64)
    hi, ms_hi = ctx.gather_rate(200_000, 96, 8, batches=64)
    assert 0 < lo < 1e5 and 0 < hi < 1e5 and ms_lo > 0 and ms_hi > 0
    assert hi > 0.8 * lo
    g128, _ = ctx.gather_rate(50_000, 128, 4, batches=16)
    assert g128 > 0
    for bad in ((1000, 300, 4), (1000, 96, 0), (1000, 96, 17), (0, 96, 4)):
        with pytest.raises(WvError):
            ctx.gather_rate(*bad, batches=4)


def test_line_rate(ctx):
    """The random-line measurement reports a positive, finite load rate; an
    array that fits the L2 caches is served at least as fast as one of 1 GiB;
    sizes outside its range are refused."""
    from weavess_amd import WvError
    small, ms_s = ctx.line_rate(4 << 20, iters=8)
    big, ms_b = ctx.line_rate(1 << 30, iters=8)
    assert 0 < big < 1e4 and 0 < small < 1e4 and ms_s > 0 and ms_b > 0
    assert small > 0.8 * big
    for nbytes, iters in (((1 << 20) - 16, 4), ((1 << 20) + 8, 4), (1 << 20, 0)):
        with pytest.raises(WvError):
            ctx.line_rate(nbytes, iters=iters)
