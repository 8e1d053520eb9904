"""Pins of the oracle's DPG refine (src/component_refine.cpp:728-790,
component_prune.cpp:124-186, component_conn.cpp:228-268)."""
import numpy as np

import oracle as O


def test_dpg_points_on_a_line():
    """Node n at x = n: its list is n-1, n+1, n-2, n+2, ...; an entry at
    distance k is occluded exactly by the k-1 nearer ones on its side, so
    the hits sort as 0,0,1,1,... and the cut is range // 2: an interior
    node keeps n-1, n+1, ..., its first `range` entries.  Every edge is then
    mirrored."""
    N, rng_ = 40, 4
    base = np.arange(N, dtype=np.float32)[:, None]
    K = 2 * rng_
    ids = np.zeros((N, K), np.uint32)
    deg = np.zeros(N, np.uint32)
    for n in range(N):
        nb = sorted((j for j in range(N) if j != n), key=lambda j: ((j - n) ** 2, j))[:K]
        ids[n] = nb
        deg[n] = K
    g, gd, gdeg = O.dpg(base, ids, deg, rng_)
    pruned = {}
    for n in range(rng_, N - rng_):
        pruned[n] = [n - 1, n + 1, n - 2, n + 2]
        row = set(g[n, :gdeg[n]].tolist())
        assert set(pruned[n]) <= row
    for n in range(N):
        for m in g[n, :gdeg[n]].tolist():
            assert n in g[m, :gdeg[m]].tolist()  # mirrored


def _dpg_pruned(base, ids, deg, rng_):
    x = base.astype(np.float64)
    out = []
    for n in range(len(base)):
        p = sorted({int(v) for v in ids[n, :deg[n]] if v != n}, key=lambda j: (((x[n] - x[j]) ** 2).sum(), j))
        L = min(len(p), 2 * rng_)
        p = p[:L]
        if L <= rng_:
            out.append(p)
            continue
        dn = [((x[n] - x[j]) ** 2).sum() for j in p]
        hit = [sum(((x[p[i]] - x[p[j]]) ** 2).sum() < dn[j] for i in range(j)) for j in range(L)]
        cut = sorted(hit)[rng_]
        out.append([p[j] for j in range(L) if hit[j] <= cut][:rng_])
    return out


def test_dpg_rows_are_selection_plus_reverse(small_int):
    base, _ = small_int
    base = base[:400]
    kid, _, kdeg = O.nndescent(base, 20, 24, 8, 24, 6, seed=1)
    g, gd, gdeg = O.dpg(base, kid, kdeg, 10)
    sel = _dpg_pruned(base, kid, kdeg, 10)
    x = base.astype(np.float64)
    for n in range(len(base)):
        want = set(sel[n]) | {m for m in range(len(base)) if n in sel[m]}
        row = g[n, :gdeg[n]].tolist()
        assert set(row) == want and len(row) == len(want), n
        keys = [(float(gd[n, i]), row[i]) for i in range(len(row))]
        assert keys == sorted(keys)
        assert np.allclose(gd[n, :gdeg[n]], ((x[n] - x[row]) ** 2).sum(1))
