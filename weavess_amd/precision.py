"""Host emulation of the local join's low-precision distance formula.

The join computes d(a, b) = |a|^2 + |b|^2 - 2 a.b on MFMA from quantised
rows: BF16 (round to nearest even), or FP8 E4M3 with a per-row absmax/448
scale; products of BF16 or FP8 values are exact in FP32 and the sums are
FP32, as on the matrix cores.  These helpers reproduce that arithmetic on
the host to state the join's error against FP64 on the source rows.  They
measure; the device never calls them.
"""
from __future__ import annotations

import numpy as np


def bf16_round(x):
    """float32 -> bfloat16 -> float32, round to nearest even (the device's
    conversion in the join's row preparation)."""
    b = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32)


def e4m3_round(x):
    """Round to OCP FP8 E4M3 (3 mantissa bits, exponent bias 7, subnormal
    step 2^-9), nearest even; |x| <= 448 already (the join clamps)."""
    x = np.asarray(x, np.float64)
    ax = np.abs(x)
    e = np.floor(np.log2(np.maximum(ax, 2.0 ** -6)))
    step = 2.0 ** (e - 3)
    return np.sign(x) * np.round(ax / step) * step


def quantise(X, dtype):
    """Rows as the join sees them, in FP64."""
    if dtype == "bf16":
        return bf16_round(X).astype(np.float64)
    if dtype == "fp8":
        sc = np.abs(X).max(1, keepdims=True) / 448.0
        sc[sc == 0] = 1.0
        return e4m3_round(np.clip(X / sc, -448, 448)) * sc
    return np.asarray(X, np.float64)


def join_formula_rel_err(base, ids, deg, dtype, nodes):
    """Max and mean relative error of the join's formula on the graph edges
    of `nodes` (the near pairs, where the formula cancels most) against FP64
    on the source rows; returns (max, mean, pairs with nonzero distance)."""
    nodes = np.asarray(nodes)
    a_idx = np.repeat(nodes, [int(deg[n]) for n in nodes])
    b_idx = np.concatenate([ids[n, :deg[n]] for n in nodes]).astype(np.int64)
    A, B = base[a_idx].astype(np.float64), base[b_idx].astype(np.float64)
    exact = ((A - B) ** 2).sum(1)
    qa = quantise(base[a_idx], dtype).astype(np.float32)
    qb = quantise(base[b_idx], dtype).astype(np.float32)
    na = (qa * qa).sum(1, dtype=np.float32)
    nb = (qb * qb).sum(1, dtype=np.float32)
    dot = (qa * qb).sum(1, dtype=np.float32)
    approx = np.maximum((na + nb) - np.float32(2.0) * dot, np.float32(0.0)).astype(np.float64)
    ok = exact > 0
    if not ok.any():
        return 0.0, 0.0, 0
    rel = np.abs(approx[ok] - exact[ok]) / exact[ok]
    return float(rel.max()), float(rel.mean()), int(ok.sum())
