"""Vamana passes on the GPU vs the CPU oracle (build-order distances on
both sides: equal bit for bit on integer and continuous data)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _ep(base):
    x = base.astype(np.float64)
    return int(((x - x.mean(0)) ** 2).sum(1).argmin())


def _same(g, o):
    gi, gd, gdeg = g
    oi, od, odeg = o
    assert np.array_equal(gdeg, odeg)
    for i in range(len(gdeg)):
        assert gi[i, :gdeg[i]].tolist() == oi[i, :odeg[i]].tolist(), i
        assert np.array_equal(gd[i, :gdeg[i]], od[i, :odeg[i]]), i


@pytest.mark.parametrize("data,R,L,C", [("small_int", 24, 40, 300), ("small_float", 16, 32, 200)])
def test_vamana_two_passes_bit_exact(ctx, request, data, R, L, C):
    """INIT_RAND's random graph, then alpha 1 and alpha 2 passes."""
    from weavess_amd import Index
    base, _ = request.getfixturevalue(data)
    rid, _, rcnt = O.init_random(base, R, seed=1)
    ep = _ep(base)
    ix = Index(ctx, base)
    ix.set_graph(rid, rcnt)
    prev = (rid, rcnt)
    for alpha in (1.0, 2.0):
        print("vamana", data, R, L, C, alpha, flush=True)
        st = ix.build_vamana(L, R, C, ep, alpha, seed=2)
        g = ix.get_graph()
        o = O.vamana(base, prev[0], prev[1], L, R, C, alpha, ep, seed=2)
        _same(g, o)
        assert st["conn_edges"] == 0
        prev = (o[0], o[2])
    ix.close()


def test_vamana_all_candidates(ctx):
    from weavess_amd import Index
    N = 30
    base = np.random.default_rng(3).integers(0, 12, size=(N, 5)).astype(np.float32)
    kid = np.array([[j for j in range(N) if j != i] for i in range(N)], np.uint32)
    kdeg = np.full(N, N - 1, np.uint32)
    ix = Index(ctx, base)
    ix.set_graph(kid, kdeg)
    ix.build_vamana(N - 1, 4, N + 5, 0, 1.2)
    g = ix.get_graph()
    ix.close()
    _same(g, O.vamana(base, kid, kdeg, N - 1, 4, N + 5, 1.2, 0))


def test_vamana_rejects_bad_params(ctx, small_float):
    from weavess_amd import Index, WvError
    base, _ = small_float
    ix = Index(ctx, base[:100])
    rid, _, rcnt = O.init_random(base[:100], 8, seed=1)
    ix.set_graph(rid, rcnt)
    for bad in [dict(alpha=0.5), dict(alpha=float("nan")), dict(R=0), dict(L=100), dict(ep=100)]:
        p = dict(L=20, R=8, C=50, ep=0, alpha=1.0)
        p.update(bad)
        with pytest.raises(WvError):
            ix.build_vamana(p["L"], p["R"], p["C"], p["ep"], p["alpha"])
    ix.close()
