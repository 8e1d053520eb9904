"""CPU pins of the learned-termination oracle (FP64 gradient, BF16 logits, features)."""
import math

import numpy as np
import pytest

import oracle as O


def _model(seed=0, std=True):
    g = np.random.default_rng(seed)
    m = O.TermModel()
    p = np.concatenate([g.normal(0, 0.5, 256), g.normal(0, 0.1, 32), g.normal(0, 0.3, 32),
                        g.normal(0, 0.1, 1)]).astype(np.float32)
    import ctypes
    ctypes.memmove(ctypes.addressof(m), p.ctypes.data, 321 * 4)
    for i in range(8):
        m.mean[i] = float(g.normal(0, 0.2)) if std else 0.0
        m.inv_std[i] = float(g.uniform(0.5, 2.0)) if std else 1.0
    return m, p


def _loss(m, X, y, pw):
    return O.mlp_grad(m, X, y, pw)[1]


def test_mlp_grad_matches_finite_differences():
    import ctypes
    g = np.random.default_rng(3)
    X = g.normal(size=(200, 8)).astype(np.float32)
    y = (g.uniform(size=200) < 0.4).astype(np.float32)
    m, p = _model(1)
    grad, loss = O.mlp_grad(m, X, y, 2.0)
    base = np.frombuffer(bytes(m), np.float32, 321).copy()
    for idx in list(range(0, 256, 37)) + [256, 270, 288, 300, 320]:
        h = 2.0 ** -9
        pp = base.copy(); pp[idx] += h
        ctypes.memmove(ctypes.addressof(m), pp.ctypes.data, 321 * 4)
        lp = _loss(m, X, y, 2.0)
        pm = base.copy(); pm[idx] -= h
        ctypes.memmove(ctypes.addressof(m), pm.ctypes.data, 321 * 4)
        lm = _loss(m, X, y, 2.0)
        ctypes.memmove(ctypes.addressof(m), base.ctypes.data, 321 * 4)
        fd = (lp - lm) / (float(pp[idx]) - float(pm[idx]))
        assert abs(fd - grad[idx]) <= 1e-4 * max(1.0, abs(grad[idx])), (idx, fd, grad[idx])


def test_mlp_logit_bf16_close_to_fp64():
    g = np.random.default_rng(4)
    X = g.normal(size=(500, 8)).astype(np.float32)
    m, _ = _model(2)
    zb = O.mlp_logit(m, X, bf16=True)
    zf = O.mlp_logit(m, X, bf16=False)
    # BF16 has 8 significant bits: the logit error stays within a few percent of its scale
    scale = np.abs(zf).mean() + 1.0
    assert np.max(np.abs(zb - zf)) < 0.05 * scale
    assert np.corrcoef(zb, zf)[0, 1] > 0.999


def test_term_features_hand_worked():
    f = O.term_features(hops=3, kth=8.0, best=2.0, entry=4.0, kth4=12.0, stale=7, L=64, k=10)
    assert f[0] == 2.0 and f[1] == 2.0 and f[2] == 0.5 and f[3] == 2.0
    assert f[4] == 1.0 and f[5] == 6.0 and f[6] == pytest.approx(math.log2(10)) and f[7] == 3.0
    # a zero entry distance is clamped, not infinite
    z = O.term_features(hops=0, kth=0.0, best=0.0, entry=0.0, kth4=0.0, stale=0, L=1, k=1)
    assert np.all(np.isfinite(z))


def test_oracle_traces_and_settle(small_float, small_float_graph):
    base, q = small_float
    g, _, deg = small_float_graph
    ids, dist, ev, hp, trace, tlen, settle = O.search_greedy_ext(g, deg, base, q, 48, 10, seed=5, trace_cap=200)
    ref = O.search_greedy(g, deg, base, q, 48, 10, seed=5)
    assert np.array_equal(ids, ref[0]) and np.array_equal(hp, ref[3])
    assert np.all(tlen == np.minimum(hp, 200)) and np.all(settle <= hp)
    # the first feature is log2(1 + expansions)
    assert np.allclose(trace[0, :3, 0], np.log2(2 + np.arange(3)))
    # stopping at the settle expansion reproduces the full top-k: a model whose
    # logit is +inf past settle would be exact; check with an always-stop
    # model, evaluated after every 4th expansion: it stops at the 4th (or the
    # search ends first)
    m = O.TermModel()
    m.b2 = 100.0
    m.logit_threshold = 0.0
    m.min_hops = 1
    for i in range(8):
        m.inv_std[i] = 1.0
    ids1, _, _, hp1, *_ = O.search_greedy_ext(g, deg, base, q, 48, 10, seed=5, model=m)
    assert np.all(hp1 == np.minimum(hp, 4)) and np.any(hp > 4)


def test_mlp_loss_closed_form():
    """All weights zero: the logit is b2 for every sample, so the weighted
    BCE is the mean of w * softplus(-b2) (y = 1, w = pos_weight) and
    softplus(b2) (y = 0), and only db2 is non-zero: mean(w * (sigmoid(b2) - y))."""
    m = O.TermModel()
    for i in range(8):
        m.inv_std[i] = 1.0
    m.b2 = 0.75  # exact in float32
    X = np.random.default_rng(5).normal(size=(40, 8)).astype(np.float32)
    y = np.array([1.0] * 15 + [0.0] * 25, np.float32)
    grad, loss = O.mlp_grad(m, X, y, 3.0)
    sp = lambda z: math.log1p(math.exp(z))
    want = (15 * 3.0 * sp(-0.75) + 25 * sp(0.75)) / 40
    assert loss == pytest.approx(want, rel=1e-12)
    sig = 1 / (1 + math.exp(-0.75))
    assert grad[320] == pytest.approx((15 * 3.0 * (sig - 1) + 25 * sig) / 40, rel=1e-12)
    assert np.all(grad[:320] == 0.0)


def test_mlp_logit_closed_form():
    """One hidden unit h = relu((f0 - 1) * 2) with w2 = 1.5 and b2 = -1:
    z = 1.5 * h - 1, in FP64 and in the BF16 order (inputs chosen so every
    product is exact in BF16)."""
    m = O.TermModel()
    m.mean[0] = 1.0
    m.inv_std[0] = 2.0
    m.w1[0][0] = 1.0
    m.w2[0] = 1.5
    m.b2 = -1.0
    f0 = np.array([0.0, 1.0, 1.25, 2.0, 3.5], np.float32)
    X = np.zeros((5, 8), np.float32)
    X[:, 0] = f0
    want = 1.5 * np.maximum((f0 - 1.0) * 2.0, 0.0) - 1.0
    assert np.array_equal(O.mlp_logit(m, X, bf16=False), want)
    assert np.array_equal(O.mlp_logit(m, X, bf16=True), want)


def _py_search_features(graph, deg, base, q, entries, L, k):
    """Sequential RouteGreedy (src/component_route.cpp:16-47) in plain Python
    with exact integer distances, recording the state features after every
    expansion and the settle point, for tiny inputs."""
    import bisect
    dist = lambda i: int(((base[i].astype(np.int64) - q.astype(np.int64)) ** 2).sum())
    pool = sorted((dist(e), int(e)) for e in entries)
    flags = [True] * L
    seen = set()
    kidx = min(k, L) - 1
    entry_d = pool[0][0]
    hist = [pool[kidx][0]] * 4
    hops = settle = 0
    rows = []
    kk = 0
    while kk < L:
        nk = L
        if flags[kk]:
            flags[kk] = False
            hops += 1
            u = pool[kk][1]
            for m in range(int(deg[u])):
                nb = int(graph[u, m])
                if nb in seen:
                    continue
                seen.add(nb)
                key = (dist(nb), nb)
                if key >= pool[L - 1]:
                    continue
                pos = bisect.bisect_left(pool, key)
                if pool[pos] == key:
                    r = L + 1
                else:
                    pool.insert(pos, key)
                    flags.insert(pos, True)
                    pool.pop()
                    flags.pop()
                    r = pos
                nk = min(nk, r)
                if r <= kidx:
                    settle = hops
            kth, best = pool[kidx][0], pool[0][0]
            rows.append([math.log2(1 + hops), kth / entry_d, best / entry_d, math.log2(entry_d),
                         (hist[3] - kth) / entry_d, math.log2(L), math.log2(k), math.log2(1 + hops - settle)])
            hist = [kth] + hist[:3]
        kk = nk if nk <= kk else kk + 1
    return [p[1] for p in pool[:k]], np.array(rows), settle, hops


@pytest.mark.parametrize("L,k", [(8, 1), (12, 5), (20, 5)])
def test_search_features_brute_force(L, k):
    """Per-expansion features, settle points and results of the oracle's
    traced search equal a plain-Python sequential search on integer data
    (many equal distances, ordered by id)."""
    g = np.random.default_rng(31)
    N, d = 300, 8
    base = g.integers(0, 16, (N, d)).astype(np.float32)
    qs = g.integers(0, 16, (6, d)).astype(np.float32)
    deg = g.integers(3, 7, N).astype(np.uint32)
    graph = np.full((N, 6), 0xFFFFFFFF, np.uint32)
    for i in range(N):
        graph[i, :deg[i]] = g.choice(N, deg[i], replace=False)
    ent = np.stack([g.choice(N, L, replace=False) for _ in qs]).astype(np.uint32)
    ids, _, ev, hp, trace, tlen, settle = O.search_greedy_ext(graph, deg, base, qs, L, k, entries=ent,
                                                              trace_cap=400)
    for i in range(len(qs)):
        pids, rows, pst, php = _py_search_features(graph, deg, base, qs[i], ent[i], L, k)
        assert ids[i].tolist() == pids and hp[i] == php and settle[i] == pst
        assert tlen[i] == len(rows)
        np.testing.assert_allclose(trace[i, :tlen[i]], rows, rtol=1e-6, atol=1e-6)
