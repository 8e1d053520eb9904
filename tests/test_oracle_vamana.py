"""Pins of the oracle's Vamana pass (src/component_refine.cpp:794-958,
component_prune.cpp:245-297)."""
import numpy as np
import pytest

import oracle as O


def _pts(N, d, seed):
    return np.random.default_rng(seed).integers(0, 12, size=(N, d)).astype(np.float32)


def _d2(base):
    x = base.astype(np.float64)
    return ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)


def _complete(N):
    ids = np.array([[j for j in range(N) if j != i] for i in range(N)], np.uint32)
    return ids, np.full(N, N - 1, np.uint32)


def _robust(D, n, cand, R, alpha):
    """Vamana's prune (DiskANN RobustPrune with the fill): by (distance, id),
    keep p unless a kept r has alpha * d(r, p) < d(n, p), at most R; then
    the nearest dropped ones up to R."""
    cand = sorted((p for p in cand if p != n), key=lambda j: (D[n, j], j))
    keep, drop = [], []
    for p in cand:
        if len(keep) == R:
            break
        if any(np.float32(alpha) * np.float32(D[r, p]) < np.float32(D[n, p]) for r in keep):
            drop.append(p)
        else:
            keep.append(p)
    return keep + drop[:R - len(keep)]


def test_vamana_huge_alpha_gives_knn_graph():
    """No occlusion: every list is the R nearest candidates; with every node
    a candidate the pass returns the exact R-NN graph (as sets)."""
    N, R = 30, 6
    base = _pts(N, 5, 0)
    D = _d2(base)
    ids, deg = _complete(N)
    g, gd, gdeg = O.vamana(base, ids, deg, N - 1, R, N + 5, 1e30, 2)
    for n in range(N):
        want = set(sorted((j for j in range(N) if j != n), key=lambda j: (D[n, j], j))[:R])
        assert set(g[n, :gdeg[n]].tolist()) == want


@pytest.mark.parametrize("alpha", [1.0, 1.2, 2.0])
def test_vamana_all_candidates_matches_robust_prune(alpha):
    N, R = 30, 4
    base = _pts(N, 5, 3)
    D = _d2(base)
    ids, deg = _complete(N)
    g, gd, gdeg = O.vamana(base, ids, deg, N - 1, R, N + 5, alpha, 0)
    own = [_robust(D, n, range(N), R, alpha) for n in range(N)]
    for t in range(N):
        inc = [s for s in range(N) if t in own[s] and s not in own[t]]
        temp = own[t] + inc
        want = temp if len(temp) <= R else _robust(D, t, temp, R, alpha)
        assert g[t, :gdeg[t]].tolist() == want, t
        assert np.allclose(gd[t, :gdeg[t]], D[t, want])


def test_vamana_two_passes_from_random_graph(small_int):
    """INIT_RAND then alpha 1 and alpha 2 passes: lists of at most R
    distinct ids, no self loops, and a searchable graph."""
    base, q = small_int
    N, R = len(base), 24
    rid, rdist, rcnt = O.init_random(base, R, seed=1)
    x = base.astype(np.float64)
    ep = int(((x - x.mean(0)) ** 2).sum(1).argmin())
    g1, _, d1 = O.vamana(base, rid, rcnt, 40, R, 300, 1.0, ep, seed=2)
    g2, _, d2 = O.vamana(base, g1, d1, 40, R, 300, 2.0, ep, seed=2)
    for g, dg in ((g1, d1), (g2, d2)):
        assert dg.max() <= R
        for i in range(N):
            row = g[i, :dg[i]].tolist()
            assert i not in row and len(set(row)) == len(row)
    gt, _ = O.exact_knn(base, q, 10)
    from weavess_amd.entries import centroid_entries
    # alpha = 2 on squared distances occludes little: pass 2's lists are
    # close to the 24 nearest, with fewer long edges between clusters
    for g, dg, L in ((g1, d1, 20), (g2, d2, 80)):
        ent = centroid_entries(ep, g, dg, len(q), L, N)
        res, *_ = O.search_greedy(g, dg, base, q, L, 10, entries=ent)
        assert O.recall(res, gt, 10) > 0.95
