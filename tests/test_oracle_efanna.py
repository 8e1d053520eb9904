"""Pins of the oracle's EFANNA refine (src/component_refine.cpp:959-1183)."""
import numpy as np
import pytest

import oracle as O


def _int_data(N, d, seed):
    return np.random.default_rng(seed).integers(0, 16, size=(N, d)).astype(np.float32)


def _brute_rows(base, rows, K):
    """Per row: ids below N, self and repeats dropped, sorted by (dist, id)
    with FP64 distances (exact on small integers), first K."""
    out = []
    for i, r in enumerate(rows):
        ids = sorted({int(x) for x in r if x < len(base) and x != i})
        dd = [float(((base[i].astype(np.float64) - base[j]) ** 2).sum()) for j in ids]
        out.append([j for _, j in sorted(zip(dd, ids))][:K])
    return out


def test_efanna_iter0_returns_the_initial_rows():
    """ITER=0: no join, no update; the pool is the initial row sorted with
    self, out-of-range and repeated ids dropped."""
    N, K, L = 60, 6, 12
    base = _int_data(N, 8, 0)
    rng = np.random.default_rng(1)
    rows = rng.integers(0, N, size=(N, L)).astype(np.uint32)
    rows[::3, 0] = np.arange(0, N, 3)        # self
    rows[::4, 1] = rows[::4, 2]              # repeats
    rows[::5, 3] = 0xFFFFFFF0                # out of range
    cnt = np.full(N, L, np.uint32)
    cnt[7] = 2
    ids, dist, deg = O.nndescent_efanna(base, K, L, 3, 8, 0, init=(rows, cnt))
    want = _brute_rows(base, [rows[i, :cnt[i]] for i in range(N)], K)
    for i in range(N):
        assert deg[i] == len(want[i]) and ids[i, :deg[i]].tolist() == want[i], i
        assert np.all(ids[i, deg[i]:] == 0xFFFFFFFF)


@pytest.mark.parametrize("seed", [0, 5])
def test_efanna_first_join_covers_the_samples(seed):
    """2S = N-1 samples per node cover every pair, and with exact insertion
    into full pools one round gives the exact K-NN graph.  KGraph's refine,
    whose first join is empty, does not."""
    N, S, K, L = 21, 10, 5, 8
    base = _int_data(N, 6, seed + 10)
    rng = np.random.default_rng(seed)
    rows = np.stack([rng.choice(np.delete(np.arange(N), i), L, replace=False) for i in range(N)]).astype(np.uint32)
    ids, _, deg = O.nndescent_efanna(base, K, L, S, 8, 1, seed=seed, init=(rows, None))
    want = _brute_rows(base, [np.delete(np.arange(N), i) for i in range(N)], K)
    assert np.all(deg == K)
    assert [ids[i].tolist() for i in range(N)] == want
    d0 = np.array([[((base[i].astype(np.float64) - base[j]) ** 2).sum() for j in rows[i]] for i in range(N)],
                  np.float32)
    kg, _, _ = O.nndescent(base, K, L, L, 8, 1, seed=seed, init=(rows, d0, np.full(N, L, np.uint32)))
    assert [kg[i].tolist() for i in range(N)] != want


def test_efanna_random_init_and_recall(small_int):
    """Without an initial graph the pools start from InitRandom's; the first
    round evaluates N * C(2S, 2) pairs, and the graph is at least as good as
    KGraph's refine after the same rounds."""
    base, _ = small_int
    N = len(base)
    K, L, S, R, it = 16, 24, 8, 24, 5
    ids, _, deg, st = O.nndescent_efanna(base, K, L, S, R, it, seed=2, cand_slots=128, with_stats=True)
    assert st[0][0] == N * (2 * S) * (2 * S - 1) // 2
    gt, _ = O.exact_knn(base, base, 10, self_offset=0)
    kg, _, _ = O.nndescent(base, K, L, S, R, it, seed=2, cand_slots=128)
    assert O.recall(ids, gt, 10) >= O.recall(kg, gt, 10) >= 0.9


def test_efanna_bad_arguments():
    base = _int_data(10, 4, 0)
    with pytest.raises(ValueError):
        O.nndescent_efanna(base, 4, 8, 5, 8, 1)          # 2S >= N
    with pytest.raises(ValueError):
        O.nndescent_efanna(base, 4, 8, 2, 8, 1, init=(np.zeros((10, 9), np.uint32), None))  # stride > L
