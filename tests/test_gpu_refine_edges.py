"""Refiner edge cases on the GPU vs the CPU oracle: the shortest and a long
row (d = 1 and d = 700 take other gather shapes), repeated points (zero
distances: the angle rule divides by zero the same way on both sides),
and a base barely larger than the lists."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _same(g, o):
    gi, gd, gdeg = g
    oi, od, odeg = o[0], o[1], o[2]
    assert np.array_equal(gdeg, odeg)
    for i in range(len(gdeg)):
        assert gi[i, :gdeg[i]].tolist() == oi[i, :odeg[i]].tolist(), i
        assert np.array_equal(gd[i, :gdeg[i]], od[i, :odeg[i]]), i


def _data(kind):
    rng = np.random.default_rng(11)
    if kind == "d1":
        return rng.integers(0, 50, size=(300, 1)).astype(np.float32)
    if kind == "d700":
        return rng.normal(size=(300, 700)).astype(np.float32)
    if kind == "repeats":
        b = rng.integers(0, 4, size=(300, 3)).astype(np.float32)  # 64 distinct points, many repeats
        return b
    return rng.integers(0, 9, size=(40, 4)).astype(np.float32)   # "tiny"


@pytest.mark.parametrize("kind", ["d1", "d700", "repeats", "tiny"])
def test_refiners_edges(ctx, kind):
    from weavess_amd import Index
    base = _data(kind)
    N = len(base)
    K = min(16, N - 2)
    kid, _, kdeg = O.nndescent(base, K, K + 4, min(6, N // 4), 16, 4, seed=1)
    x = base.astype(np.float64)
    ep = int(((x - x.mean(0)) ** 2).sum(1).argmin())
    L = min(30, N - 1)
    print("edges", kind, N, flush=True)
    ix = Index(ctx, base)
    try:
        ix.set_graph(kid, kdeg)
        ix.build_nsg(L, 8, 60, ep, seed=3)
        _same(ix.get_graph(), O.nsg(base, kid, kdeg, L, 8, 60, ep, seed=3))
        ix.set_graph(kid, kdeg)
        st, eps = ix.build_ssg(L, 8, 60.0, 3, seed=3)
        o = O.ssg(base, kid, kdeg, L, 8, 60.0, 3, seed=3)
        assert eps.tolist() == o[3].tolist()
        _same(ix.get_graph(), o)
        ix.set_graph(kid, kdeg)
        ix.build_dpg(K // 2)
        _same(ix.get_graph(), O.dpg(base, kid, kdeg, K // 2))
        ix.set_graph(kid, kdeg)
        ix.build_vamana(L, 8, 60, ep, 1.0, seed=3)
        g1 = ix.get_graph()
        o1 = O.vamana(base, kid, kdeg, L, 8, 60, 1.0, ep, seed=3)
        _same(g1, o1)
        ix.build_vamana(L, 8, 60, ep, 2.0, seed=3)
        _same(ix.get_graph(), O.vamana(base, o1[0], o1[2], L, 8, 60, 2.0, ep, seed=3))
        st = ix.build_efanna(K, K + 4, min(6, N // 4 - 1), 16, 3, seed=2)
        g = ix.get_graph()
        _same(g, O.nndescent_efanna(base, K, K + 4, min(6, N // 4 - 1), 16, 3, seed=2, cand_slots=st["cand_slots"]))
    finally:
        ix.close()
