"""NSG refine on the GPU vs the CPU oracle.  Every distance on both sides is
the build-order FP32 distance, so the graphs must be equal bit for bit on
integer and on continuous data."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _ep(base):
    x = base.astype(np.float64)
    return int(((x - x.mean(0)) ** 2).sum(1).argmin())


def _gpu_nsg(ctx, base, kid, kdeg, L, R, C, ep, seed):
    from weavess_amd import Index
    ix = Index(ctx, base)
    try:
        ix.set_graph(kid, kdeg)
        st = ix.build_nsg(L, R, C, ep, seed=seed)
        ids, dist, deg = ix.get_graph()
    finally:
        ix.close()
    return ids, dist, deg, st


def _same(g, o):
    gi, gd, gdeg = g
    oi, od, odeg = o
    assert np.array_equal(gdeg, odeg)
    for i in range(len(gdeg)):
        assert gi[i, :gdeg[i]].tolist() == oi[i, :odeg[i]].tolist(), i
        assert np.array_equal(gd[i, :gdeg[i]], od[i, :odeg[i]]), i


@pytest.mark.parametrize("data,R,L,C", [("small_int", 16, 40, 200), ("small_float", 16, 40, 200),
                                        ("small_int", 1, 24, 100), ("small_float", 40, 64, 500)])
def test_nsg_bit_exact(ctx, request, data, R, L, C):
    base, _ = request.getfixturevalue(data)
    kid, _, kdeg = O.nndescent(base, 24, 32, 10, 32, 6, seed=1, cand_slots=128)
    ep = _ep(base)
    print("nsg", data, R, L, C, flush=True)
    gi, gd, gdeg, st = _gpu_nsg(ctx, base, kid, kdeg, L, R, C, ep, seed=2)
    oi, od, odeg, oe = O.nsg(base, kid, kdeg, L, R, C, ep, seed=2)
    print(st, "oracle conn edges", oe)
    assert st["conn_edges"] == oe and st["width"] == oi.shape[1]
    _same((gi, gd, gdeg), (oi, od, odeg))


def test_nsg_all_candidates(ctx):
    """Complete kNN graph with L = N-1 and C, R >= N: the MRNG selection plus
    reverse edges (the oracle is pinned to the brute force there)."""
    N = 30
    base = np.random.default_rng(0).integers(0, 12, size=(N, 5)).astype(np.float32)
    kid = np.array([[j for j in range(N) if j != i] for i in range(N)], np.uint32)
    kdeg = np.full(N, N - 1, np.uint32)
    gi, gd, gdeg, st = _gpu_nsg(ctx, base, kid, kdeg, N - 1, N, N + 5, 3, seed=0)
    oi, od, odeg, _ = O.nsg(base, kid, kdeg, N - 1, N, N + 5, 3, seed=0)
    _same((gi, gd, gdeg), (oi, od, odeg))


def test_nsg_search_beats_knn_graph(ctx, small_float):
    from weavess_amd import Index
    from weavess_amd.entries import centroid_entries
    base, q = small_float
    kid, _, kdeg = O.nndescent(base, 24, 32, 10, 32, 6, seed=1, cand_slots=128)
    ep = _ep(base)
    gi, _, gdeg, st = _gpu_nsg(ctx, base, kid, kdeg, 40, 24, 300, ep, seed=3)
    gt, _ = O.exact_knn(base, q, 10)
    ix = Index(ctx, base)
    ix.set_graph(gi, gdeg)
    ent = centroid_entries(ep, gi, gdeg, len(q), 16, len(base))
    r_nsg = ix.search(q, 16, 10, entries=ent)[0]
    ix.set_graph(kid, kdeg)
    r_knn = ix.search(q, 16, 10, seed=0)[0]
    ix.close()
    print("recall nsg", O.recall(r_nsg, gt, 10), "knn", O.recall(r_knn, gt, 10), st)
    assert O.recall(r_nsg, gt, 10) > O.recall(r_knn, gt, 10)


def test_nsg_rejects_bad_params(ctx, small_float):
    from weavess_amd import Index, WvError
    base, _ = small_float
    ix = Index(ctx, base)
    with pytest.raises(WvError):
        ix.build_nsg(40, 16, 200, 0)          # no graph yet
    kid, _, kdeg = O.nndescent(base[:100], 8, 16, 4, 16, 2, seed=1)
    ix.close()
    ix = Index(ctx, base[:100])
    ix.set_graph(kid, kdeg)
    for bad in [dict(L=100), dict(R=0), dict(R=300), dict(C=2000), dict(ep=100)]:
        p = dict(L=20, R=8, C=50, ep=0)
        p.update(bad)
        with pytest.raises(WvError):
            ix.build_nsg(p["L"], p["R"], p["C"], p["ep"])
    ix.close()
