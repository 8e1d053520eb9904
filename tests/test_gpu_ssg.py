"""SSG refine on the GPU vs the CPU oracle (build-order distances and the
FP32 angle rule on both sides: equal bit for bit)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _gpu(ctx, base, kid, kdeg, L, R, angle, n_try, seed):
    from weavess_amd import Index
    ix = Index(ctx, base)
    try:
        ix.set_graph(kid, kdeg)
        st, eps = ix.build_ssg(L, R, angle, n_try, seed)
        g = ix.get_graph()
    finally:
        ix.close()
    return g, st, eps


def _same(g, o):
    gi, gd, gdeg = g
    oi, od, odeg = o
    assert np.array_equal(gdeg, odeg)
    for i in range(len(gdeg)):
        assert gi[i, :gdeg[i]].tolist() == oi[i, :odeg[i]].tolist(), i
        assert np.array_equal(gd[i, :gdeg[i]], od[i, :odeg[i]]), i


@pytest.mark.parametrize("data,L,R,angle", [("small_int", 50, 20, 60.0), ("small_float", 100, 24, 60.0),
                                            ("small_int", 30, 4, 45.0)])
def test_ssg_bit_exact(ctx, request, data, L, R, angle):
    base, _ = request.getfixturevalue(data)
    kid, _, kdeg = O.nndescent(base, 24, 32, 10, 32, 6, seed=1, cand_slots=128)
    print("ssg", data, L, R, angle, flush=True)
    g, st, eps = _gpu(ctx, base, kid, kdeg, L, R, angle, 10, 2)
    oi, od, odeg, oeps, oe = O.ssg(base, kid, kdeg, L, R, angle, 10, seed=2)
    print(st, "oracle links", oe)
    assert eps.tolist() == oeps.tolist() and st["conn_edges"] == oe
    _same(g, (oi, od, odeg))


def test_ssg_two_cliques(ctx):
    rng = np.random.default_rng(5)
    A = rng.integers(0, 5, size=(12, 4)).astype(np.float32)
    B = rng.integers(0, 5, size=(9, 4)).astype(np.float32) + 40
    base = np.concatenate([A, B])
    N = len(base)
    ids = np.full((N, N), 0xFFFFFFFF, np.uint32)
    deg = np.zeros(N, np.uint32)
    for grp in (range(0, 12), range(12, N)):
        for i in grp:
            row = [j for j in grp if j != i]
            ids[i, :len(row)] = row
            deg[i] = len(row)
    for seed in range(3):
        g, st, eps = _gpu(ctx, base, ids, deg, 40, 30, 60.0, 3, seed)
        oi, od, odeg, oeps, oe = O.ssg(base, ids, deg, 40, 30, 60.0, 3, seed=seed)
        assert eps.tolist() == oeps.tolist() and st["conn_edges"] == oe
        _same(g, (oi, od, odeg))


def test_ssg_rejects_bad_params(ctx, small_float):
    from weavess_amd import Index, WvError
    base, _ = small_float
    ix = Index(ctx, base[:100])
    kid, _, kdeg = O.nndescent(base[:100], 8, 16, 4, 16, 2, seed=1)
    ix.set_graph(kid, kdeg)
    for bad in [dict(R=0), dict(L=0), dict(n_try=100), dict(angle=0.0), dict(angle=180.0)]:
        p = dict(L=20, R=8, angle=60.0, n_try=3)
        p.update(bad)
        with pytest.raises(WvError):
            ix.build_ssg(p["L"], p["R"], p["angle"], p["n_try"])
    ix.close()
