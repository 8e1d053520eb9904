"""Learned early termination of greedy search.

Workflow (all GPU work through the C ABI):

1. ``collect_traces``: plain fixed-L search on training queries, recording the
   feature vector g after every expansion and, per query, the expansion after
   which the top-k never changed again (``settle``).  A state is labelled 1
   ("stopping here loses nothing") when its expansion count is >= settle.
2. ``train``: the 8 -> 32 -> 1 MLP trained on the device (Adam, weighted BCE).
3. ``calibrate``: the lowest logit threshold whose recall@k on calibration
   queries reaches the target (the search stops when the network's logit
   exceeds the threshold).
4. Search with ``Index.search_ext(..., model=m)``: the BF16 inference step runs
   inside the search loop after every 4th expansion; plain fixed-L search remains
   available as the fallback and as the reference for the recall loss.
"""
from __future__ import annotations

import numpy as np

from ._capi import Context, Index, TermModel


def traces_to_samples(trace, tlen, settle):
    """Flatten per-query traces into (X, y): row h of query q is the state
    after h+1 expansions; label 1 when h+1 >= settle[q]."""
    rows, labels = [], []
    for q in range(trace.shape[0]):
        n = int(tlen[q])
        if n == 0:
            continue
        rows.append(trace[q, :n])
        labels.append((np.arange(1, n + 1) >= settle[q]).astype(np.float32))
    if not rows:
        return np.zeros((0, 8), np.float32), np.zeros(0, np.float32)
    return np.concatenate(rows).astype(np.float32), np.concatenate(labels)


def collect_traces(index: Index, queries, L, k, seed=0, cap=None, L_per_q=None, k_per_q=None):
    """Plain search with traces; returns X, y and the plain results."""
    if cap is None:
        cap = int(np.max(L_per_q)) * 2 if L_per_q is not None else 2 * L
    ids, dist, ev, hp, st, trace, tlen, settle = index.search_ext(
        queries, L, k, seed=seed, trace_cap=cap, L_per_q=L_per_q, k_per_q=k_per_q)
    X, y = traces_to_samples(trace, tlen, settle)
    return X, y, (ids, ev, hp, settle)


def train(ctx: Context, X, y, epochs=20, batch=4096, lr=3e-3, pos_weight=1.0, seed=0,
          min_hops=1) -> tuple:
    """Train a fresh network on (X, y); returns (model, training stats)."""
    m = TermModel.init(seed)
    st = ctx.mlp_train(m, X, y, epochs=epochs, batch=batch, lr=lr, pos_weight=pos_weight, seed=seed)
    m.min_hops = min_hops
    return m, st


def recall_at_k(res, gt, k):
    hit = (res[:, :k, None] == gt[:, None, :k]).any(axis=2)
    return float(hit.mean())


def calibrate(index: Index, queries, gt, model: TermModel, L, k, target, seed=0,
              thresholds=None, L_per_q=None, k_per_q=None, bisect=False):
    """Lowest logit threshold with recall@k >= target on (queries, gt).

    Scans from the most conservative threshold down and stops at the first
    miss; with ``bisect`` it binary-searches the sorted thresholds instead
    (recall is taken as non-decreasing in the threshold: a higher threshold
    only stops later).  Returns (threshold, recall, mean evaluations) or None
    if even the most conservative threshold misses the target."""
    if thresholds is None:
        thresholds = np.arange(-6.0, 8.01, 0.5)
    ts = sorted(float(t) for t in thresholds)
    saved = model.logit_threshold

    def trial(t):
        model.logit_threshold = t
        ids, _, ev, hp, st, *_ = index.search_ext(queries, L, k, seed=seed, model=model,
                                                  L_per_q=L_per_q, k_per_q=k_per_q)
        r = recall_at_k(ids, gt, min(k, gt.shape[1])) if L_per_q is None else None
        return (t, r, float(ev.mean())) if r is not None and r >= target else None

    best = None
    if bisect:
        lo, hi = -1, len(ts) - 1  # ts[hi] is the candidate; everything <= lo misses
        best = trial(ts[hi])
        if best is not None:
            while hi - lo > 1:
                mid = (lo + hi) // 2
                got = trial(ts[mid])
                if got is not None:
                    hi, best = mid, got
                else:
                    lo = mid
    else:
        for t in reversed(ts):  # conservative -> aggressive
            got = trial(t)
            if got is None:
                break
            best = got
    model.logit_threshold = saved if best is None else best[0]
    return best
