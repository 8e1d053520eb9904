"""Pins of the oracle's SSG refine (src/component_refine.cpp:550-727,
component_candidate.cpp:89-108, component_prune.cpp:71-122,
component_conn.cpp:169-226)."""
import numpy as np

import oracle as O


def test_ssg_threshold_is_cos_a():
    assert O.ssg_threshold(60.0) == np.float32(np.cos(np.float64(np.float32(60.0) / 180) * np.pi))
    assert abs(O.ssg_threshold(60.0) - 0.5) < 1e-7 and abs(O.ssg_threshold(90.0)) < 1e-7


def test_ssg_angle_prune_by_hand():
    """Node 0 at the origin with satellites at angles 0, 30, 90 and 200
    degrees, radii 1..4: 0 is kept, 30 is within 60 degrees of it and is
    dropped, 90 and 200 are kept.  The satellites list only node 0, so the
    two-hop candidates add nothing."""
    ang = np.radians([0.0, 30.0, 90.0, 200.0])
    rad = np.array([1.0, 2.0, 3.0, 4.0])
    pts = np.stack([rad * np.cos(ang), rad * np.sin(ang)], 1)
    base = np.concatenate([np.zeros((1, 2)), pts]).astype(np.float32)
    N = len(base)
    ids = np.zeros((N, 4), np.uint32)
    deg = np.ones(N, np.uint32)
    ids[0] = [1, 2, 3, 4]
    deg[0] = 4
    g, gd, gdeg, eps, e = O.ssg(base, ids, deg, 10, 8, 60.0, 1, seed=0)
    assert g[0, :3].tolist() == [1, 3, 4]  # then any links of the connectivity pass
    assert 2 not in g[0, :gdeg[0]].tolist() or gdeg[0] > 3


def _reach(g, dg, r):
    seen, st = {r}, [r]
    while st:
        u = st.pop()
        for v in g[u, :dg[u]].tolist():
            if v not in seen:
                seen.add(v)
                st.append(v)
    return seen


def test_ssg_conn_two_cliques():
    """Cliques A = 0..11 and B = 12..20 with lists only inside: a root in A
    links min(A) -> min(B); a root in B reaches only B and links
    min(B) -> min(A).  One edge if every root is on one side, else two."""
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
    for seed in range(6):
        g, gd, gdeg, eps, e = O.ssg(base, ids, deg, 40, 30, 60.0, 3, seed=seed)
        sides = {int(r) < 12 for r in eps}
        assert e == len(sides)
        first_a = int(eps[0]) < 12
        a0, b0 = (0, 12)
        assert (b0 in g[a0, :gdeg[a0]].tolist()) == (first_a or len(sides) == 2)
        assert (a0 in g[b0, :gdeg[b0]].tolist()) == ((not first_a) or len(sides) == 2)
        for r in eps:
            assert len(_reach(g, gdeg, int(r))) == N


def test_ssg_real_build(small_int):
    base, q = small_int
    N = len(base)
    kid, _, kdeg = O.nndescent(base, 24, 32, 10, 32, 6, seed=1, cand_slots=128)
    g, gd, gdeg, eps, e = O.ssg(base, kid, kdeg, 50, 20, 60.0, 10, seed=2)
    assert gdeg.max() <= 20 and len(set(eps.tolist())) == 10
    for i in range(N):
        row = g[i, :gdeg[i]].tolist()
        assert i not in row and len(set(row)) == len(row)
    for r in eps:
        assert len(_reach(g, gdeg, int(r))) == N
    gt, _ = O.exact_knn(base, q, 10)
    from weavess_amd.entries import subcentroid_entries
    ent = subcentroid_entries(eps, len(q), 40, N)
    res, *_ = O.search_greedy(g, gdeg, base, q, 40, 10, entries=ent)
    assert O.recall(res, gt, 10) > 0.95
