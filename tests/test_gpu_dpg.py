"""DPG refine on the GPU vs the CPU oracle (build-order distances: equal bit
for bit on integer and continuous data)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _same(g, o):
    gi, gd, gdeg = g
    oi, od, odeg = o
    assert np.array_equal(gdeg, odeg)
    assert gi.shape[1] == oi.shape[1]
    for i in range(len(gdeg)):
        assert gi[i, :gdeg[i]].tolist() == oi[i, :odeg[i]].tolist(), i
        assert np.array_equal(gd[i, :gdeg[i]], od[i, :odeg[i]]), i


@pytest.mark.parametrize("data,K,rng_", [("small_int", 24, 12), ("small_float", 32, 16), ("small_int", 24, 30)])
def test_dpg_bit_exact(ctx, request, data, K, rng_):
    from weavess_amd import Index
    base, _ = request.getfixturevalue(data)
    kid, _, kdeg = O.nndescent(base, K, K + 8, 10, 32, 6, seed=1, cand_slots=128)
    print("dpg", data, K, rng_, flush=True)
    ix = Index(ctx, base)
    ix.set_graph(kid, kdeg)
    st = ix.build_dpg(rng_)
    g = ix.get_graph()
    ix.close()
    o = O.dpg(base, kid, kdeg, rng_)
    print(st)
    assert st["width"] == o[0].shape[1]
    _same(g, o)


def test_dpg_line(ctx):
    from weavess_amd import Index
    N, rng_ = 40, 4
    base = np.arange(N, dtype=np.float32)[:, None]
    ids = np.array([sorted((j for j in range(N) if j != n), key=lambda j: ((j - n) ** 2, j))[:2 * rng_]
                    for n in range(N)], np.uint32)
    deg = np.full(N, 2 * rng_, np.uint32)
    ix = Index(ctx, base)
    ix.set_graph(ids, deg)
    ix.build_dpg(rng_)
    g = ix.get_graph()
    ix.close()
    _same(g, O.dpg(base, ids, deg, rng_))


def test_dpg_rejects_bad_params(ctx, small_float):
    from weavess_amd import Index, WvError
    base, _ = small_float
    ix = Index(ctx, base[:100])
    with pytest.raises(WvError):
        ix.build_dpg(8)          # no graph yet
    kid, _, kdeg = O.nndescent(base[:100], 8, 16, 4, 16, 2, seed=1)
    ix.set_graph(kid, kdeg)
    for bad in (0, 300):
        with pytest.raises(WvError):
            ix.build_dpg(bad)
    ix.close()


def _hub_graph(N, rng_, seed):
    """Node 0 sits at the origin and every other node lists it among its
    2*range neighbours (the rest random): it is nearly always the nearest
    entry, kept with no hits, so node 0's mirrored row holds most of the
    nodes."""
    rng = np.random.default_rng(seed)
    base = rng.normal(size=(N, 8)).astype(np.float32)
    base[0] = 0
    ids = np.empty((N, 2 * rng_), np.uint32)
    ids[0] = rng.choice(np.arange(1, N), 2 * rng_, replace=False)
    for n in range(1, N):
        others = np.concatenate([np.arange(1, n), np.arange(n + 1, N)])
        ids[n, 0] = 0
        ids[n, 1:] = rng.choice(others, 2 * rng_ - 1, replace=False)
    return base, ids, np.full(N, 2 * rng_, np.uint32)


def test_dpg_hub_row_spans_chunks(ctx):
    """A mirrored row of several hundred keys: the distinct pass runs over
    many 64-key chunks, each key compared with the one before it across the
    chunk edges, and the row is padded to the width."""
    from weavess_amd import Index
    base, ids, deg = _hub_graph(700, 4, seed=7)
    o = O.dpg(base, ids, deg, 4)
    assert o[0].shape[1] > 3 * 64
    ix = Index(ctx, base)
    ix.set_graph(ids, deg)
    st = ix.build_dpg(4)
    g = ix.get_graph()
    ix.close()
    assert st["width"] == o[0].shape[1]
    _same(g, o)
