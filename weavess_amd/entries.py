"""Centroid and sub-centroid entry points for greedy search.

`centroid_ep` mirrors ComponentRefineEntryCentroid (src/component_entry.cpp:
9-28): the navigating node ep is the base point nearest to the centroid of
the base.  The reference finds it with an approximate greedy search from a
random node; here it is the exact FP64 nearest neighbour of the centroid,
computed on the GPU.  The centroid is summed in double and rounded to float.

`centroid_entries` mirrors ComponentSearchEntryCentroid
(src/component_search_entry.cpp:33-67): per query, the first min(L, degree)
neighbours of ep, then distinct random ids until L.  `subcentroid_entries`
mirrors ComponentSearchEntrySubCentroid (src/component_search_entry.cpp:
70-99): the entry set eps first, then random ids.  The reference draws the
fill with rand() / GenRandom; here the draws come from the query's counter
stream, the same one the C++ pipeline stages use: draw t mod N,
t = 0, 1, ..., skipping ids already taken, so every set is distinct.  The
entry sets go to the search as explicit entries.
"""
from __future__ import annotations

import numpy as np

_M64 = (1 << 64) - 1
_GOLDEN = 0x9E3779B97F4A7C15
TAG_CENTROID = 0x43454E54     # "CENT"
TAG_SUBCENTROID = 0x53554243  # "SUBC"


def _fmix64(z: int) -> int:
    z ^= z >> 30
    z = (z * 0xBF58476D1CE4E5B9) & _M64
    z ^= z >> 27
    z = (z * 0x94D049BB133111EB) & _M64
    z ^= z >> 31
    return z


def _stream_of(seed: int, tag: int, idx: int) -> int:
    s = _fmix64((seed + _GOLDEN) & _M64)
    s = _fmix64(s ^ ((tag * _GOLDEN) & _M64))
    return _fmix64(s ^ ((idx + 0x632BE59BD9B4E019) & _M64))


def centroid(base: np.ndarray) -> np.ndarray:
    return (np.asarray(base, np.float64).sum(axis=0) / len(base)).astype(np.float32)


def centroid_ep(ctx, base: np.ndarray) -> int:
    """Index of the base point nearest to the centroid (exact, FP64)."""
    ids, _, _ = ctx.exact_knn(base, centroid(base)[None, :], 1)
    return int(ids[0, 0])


def head_entries(head, Q: int, L: int, N: int, seed: int, tag: int) -> np.ndarray:
    """(Q, L) uint32: the ids of `head` (below N, first occurrence, at most
    L), then the query's counter-stream fill."""
    if L > N:
        raise ValueError("L above the base size")
    first, seen = [], set()
    for nb in head:
        nb = int(nb)
        if len(first) == L:
            break
        if nb < N and nb not in seen:
            seen.add(nb)
            first.append(nb)
    out = np.empty((Q, L), np.uint32)
    for q in range(Q):
        ids = list(first)
        taken = set(seen)
        s = _stream_of(seed, tag, q)
        t = 0
        while len(ids) < L:
            v = (_fmix64((s + (t + 1) * _GOLDEN) & _M64) >> 32) % N
            t += 1
            if v in taken:
                continue
            taken.add(v)
            ids.append(v)
        out[q] = ids
    return out


def centroid_entries(ep: int, graph: np.ndarray, deg: np.ndarray, Q: int, L: int, N: int,
                     seed: int = 0) -> np.ndarray:
    """ep's neighbours, then the fill."""
    return head_entries(graph[ep, :int(deg[ep])], Q, L, N, seed, TAG_CENTROID)


def subcentroid_entries(eps, Q: int, L: int, N: int, seed: int = 0) -> np.ndarray:
    """The entry set eps (at most L ids), then the fill."""
    if len(eps) > L:
        raise ValueError("more entry points than L")
    return head_entries(eps, Q, L, N, seed, TAG_SUBCENTROID)
