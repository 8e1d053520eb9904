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
    # logit is +inf past settle would be exact; check with an always-stop model
    m = O.TermModel()
    m.b2 = 100.0
    m.logit_threshold = 0.0
    m.min_hops = 1
    for i in range(8):
        m.inv_std[i] = 1.0
    ids1, _, _, hp1, *_ = O.search_greedy_ext(g, deg, base, q, 48, 10, seed=5, model=m)
    assert np.all(hp1 == 1)
