"""Pins of the oracle's NSG refine (src/component_refine.cpp:393-540)."""
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


def _mrng(D, n, R):
    """MRNG edge selection (Fu et al., NSG): others by (distance, id); p is
    kept unless a kept r has d(r, p) < d(n, p); at most R."""
    res = []
    for p in sorted((j for j in range(len(D)) if j != n), key=lambda j: (D[n, j], j)):
        if len(res) == R:
            break
        if any(D[r, p] < D[n, p] for r in res):
            continue
        res.append(p)
    return res


def _reach(ids, deg, ep):
    seen = {ep}
    st = [ep]
    while st:
        u = st.pop()
        for v in ids[u, :deg[u]].tolist():
            if v not in seen:
                seen.add(v)
                st.append(v)
    return seen


@pytest.mark.parametrize("seed", [0, 1])
def test_nsg_all_candidates_gives_mrng_plus_reverse(seed):
    """Complete kNN graph, L = N-1, C and R >= N: every node is a candidate,
    so each pruned list is the MRNG selection; the reverse insertion appends
    the in-neighbours (ascending, not already present) without re-pruning."""
    N = 30
    base = _pts(N, 5, seed)
    D = _d2(base)
    ids, deg = _complete(N)
    ep = 3
    g, gd, gdeg, e = O.nsg(base, ids, deg, N - 1, N, N + 5, ep, seed=seed)
    own = [_mrng(D, n, N) for n in range(N)]
    for t in range(N):
        want = own[t] + [s for s in range(N) if t in own[s] and s not in own[t]]
        assert g[t, :gdeg[t]].tolist() == want, t
        assert np.allclose(gd[t, :gdeg[t]], D[t, want])
    assert e == 0  # MRNG plus its reverse edges is connected


def test_nsg_degree_cap_and_reprune():
    """R = 3: pruned lists keep the first 3 MRNG picks; a list that grows past
    R by reverse edges is re-pruned over own + incoming in key order."""
    N, R = 30, 3
    base = _pts(N, 5, 7)
    D = _d2(base)
    ids, deg = _complete(N)
    g, _, gdeg, e = O.nsg(base, ids, deg, N - 1, R, N + 5, 0, seed=3)
    own = [_mrng(D, n, R) for n in range(N)]
    rows = {}
    for t in range(N):
        inc = [s for s in range(N) if t in own[s] and s not in own[t]]
        temp = own[t] + inc
        if len(temp) <= R:
            rows[t] = temp
        else:
            cand = sorted(temp, key=lambda j: (D[t, j], j))
            res = []
            for p in cand:
                if len(res) == R:
                    break
                if not any(D[r, p] < D[t, p] for r in res):
                    res.append(p)
            rows[t] = res
    n_conn = 0
    for t in range(N):
        got = g[t, :gdeg[t]].tolist()
        assert got[:len(rows[t])] == rows[t], t
        n_conn += len(got) - len(rows[t])
    assert n_conn == e
    assert len(_reach(g, gdeg, 0)) == N


def test_nsg_conn_links_nearest_reached_node():
    """Two cliques, ep in the first: one edge, from the first clique's node
    nearest to the second clique's smallest id (every node is a candidate)."""
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
    g, gd, gdeg, e = O.nsg_conn(base, ids, deg, N - 1, N, 0, seed=1)
    assert e == 1
    D = _d2(base)
    root = min(range(12), key=lambda a: (D[a, 12], a))
    assert g[root, gdeg[root] - 1] == 12 and gd[root, gdeg[root] - 1] == D[root, 12]
    assert len(_reach(g, gdeg, 0)) == N


def test_nsg_reachable_and_clean(small_int):
    """A real build at R = 1 and R = 16: every node reachable from ep, no
    self loops or repeated ids, edges above R only from the connectivity
    pass; NSG search beats the kNN graph at small L."""
    base, q = small_int
    N = len(base)
    kid, _, kdeg = O.nndescent(base, 24, 32, 10, 32, 6, seed=1, cand_slots=128)
    x = base.astype(np.float64)
    ep = int(((x - x.mean(0)) ** 2).sum(1).argmin())
    for R in (1, 16):
        g, gd, gdeg, e = O.nsg(base, kid, kdeg, 40, R, 200, ep, seed=2)
        assert len(_reach(g, gdeg, ep)) == N
        for i in range(N):
            row = g[i, :gdeg[i]].tolist()
            assert i not in row and len(set(row)) == len(row)
        assert int(gdeg.sum()) <= N * R + e
    gt, _ = O.exact_knn(base, q, 10)
    from weavess_amd.entries import centroid_entries
    ent = centroid_entries(ep, g, gdeg, len(q), 12, N)
    r_nsg, *_ = O.search_greedy(g, gdeg, base, q, 12, 10, entries=ent)
    r_knn, *_ = O.search_greedy(kid, kdeg, base, q, 12, 10, seed=0)
    assert O.recall(r_nsg, gt, 10) > O.recall(r_knn, gt, 10)


def test_nsg_bad_arguments():
    base = _pts(10, 3, 0)
    ids, deg = _complete(10)
    with pytest.raises(ValueError):
        O.nsg(base, ids, deg, 10, 4, 20, 0)   # L >= N
    with pytest.raises(ValueError):
        O.nsg(base, ids, deg, 5, 4, 20, 10)   # ep out of range
