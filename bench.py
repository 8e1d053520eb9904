#!/usr/bin/env python3
"""Benchmark: QPS at recall@10 = 0.95, learned early termination vs plain
fixed-L greedy search, on one MI355X; NN-Descent build seconds at graph
recall 0.95.

Default workload "c5" (the configuration the metric is quoted on):
  * 2,000,000 base x 96, unit-norm, seeded 2,000-cluster Gaussian mixture
    (sigma 0.3 per dimension before normalisation); 200,000 workload
    queries plus 2,000 training and 5,000 calibration queries from the
    same mixture; k = 10 from seeded random entries.
  * Graph: NN-Descent on the GPU, K=48 (L=128, S=32, R=128, 20 rounds,
    BF16 join with FP32 rerank).
  * Plain: the smallest L whose recall@10 (exact FP64 kNN, first 20,000
    workload queries) reaches the target, by sweep and bisection; at 0.95
    (the value) and at 0.99.
  * Learned, at each target: the 8 -> 32 -> 1 stop network, trained on the
    GPU on at most 500,000 per-expansion traces of the training queries at
    pool cap 1.3 x L_plain (at most 4,096), threshold calibrated on the
    calibration queries for target + min(0.003, (1 - target) / 10), searched
    at that cap.  Beside it, the ceiling of any
    stop rule: plain search at L_plain stopped at each query's settle hop.
  * NN-Descent leg: the 1M x 128 Zipf mixture (K=64, S=16, 12 rounds) with
    L = 64, R = 128 and 128 candidate slots per node, built in FP32 and in
    BF16 (FP32 rerank); graph recall on 10,000 control nodes against exact
    FP64 kNN; the join's roofline; each dtype's build at the fewest rounds
    that reach graph recall 0.95; graph recall with 256 and 1,024 slots;
    the 1M x 128 configuration's fixed-L search (L 20-160) on the FP32 graph.
  * Precision leg: the 1M x 960 non-negative mixture (K=64 L=64 S=16 R=64,
    12 rounds) built with FP32, BF16 and FP8 joins (FP32 rerank of each
    pool's top 2K): build seconds, graph recall, the join formula's and the
    stored distances' error against FP64; and each dtype's build at the
    fewest rounds that reach graph recall 0.95.
  * Batching leg: per-query L ~ U{16..400}, k in {1, 10, 100} over the
    200,000 queries in one autotuned launch on the 2M x 96 graph.
  * CPU baseline (host cores, the oracle): plain search on a bounded sample
    of the workload queries at L_plain (ids compared with the GPU's), one
    thread too; NN-Descent at 10k x 128 with the reference's exact bounded
    insertion over 3 seeds, beside the GPU's graph recall; the 10k x 128
    search at L=100.

A "step" is one search pass over the 200,000 workload queries, with base,
graph, queries, model and outputs resident in HBM.  Each method's search
tuning is picked by the library's autotuner on 8,192 workload queries before
its timed steps (results never depend on it).  Exactly --steps steps of the
0.95 methods are timed after --warmup untimed ones (at most 2 at 0.99),
bracketed by device synchronisation.  `value` is the QPS of the faster
method among those that reach the target recall on the sample (plain
fixed-L search is the fallback the learned stop is judged against); both
are reported.  Prints ONE JSON line on stdout; progress goes to stderr.

--workload c2 / c1: plain fixed-L search only, on the 1M x 128 (K=64,
L=R=64 graph) or the 10k x 128 integer mixture (K=50, L=70, S=10, R=100).
--workload c3: NN-Descent precision on 1M x 960 (FP32 / BF16 / FP8 joins):
graph recall and distance error against FP64.
--workload nnd: the NN-Descent leg alone.
--workload c4: the batching configuration alone.
--workload graphs: the L recall 0.95 / 0.99 need on the bench's NN-Descent
graph, a denser NN-Descent build and the exact K=48 kNN graph.
--workload refine: the NSG / SSG / DPG / Vamana refiners and EFANNA.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "QPS @ recall@10=0.95 (learned vs plain); NN-Descent build s @ 0.95 graph recall"
HBM_PEAK_GBS = 8000.0
L_MAX = 4096  # largest search pool the kernel supports

WORKLOADS = {
    "c5": dict(desc="2M base x 200k queries, d=96, unit-norm 2000-cluster GMM (sigma 0.3/dim before "
                    "normalisation); NN-Descent K=48 L=128 S=32 R=128 iters=20 (BF16 join); greedy "
                    "search k=10 from random entries; learned stop vs plain fixed L",
               gen="gmm_c4", gen_kw=dict(n_query=207_000), n_work=200_000, n_train=2_000, n_cal=5_000,
               K=48, L=128, S=32, R=128, iters=20, dtype="bf16", k=10,
               sweep=[400, 800, 1600, 2400, 3200, 4096], recall_sample=20_000, control=10_000,
               learned=True),
    "c2": dict(desc="1M base x 10k queries, d=128, 1000-cluster Zipf(1.1) GMM; NN-Descent K=64 L=64 "
                    "S=16 R=64 iters=12; greedy search k=10 from random entries",
               gen="gmm_c2", gen_kw={}, K=64, L=64, S=16, R=64, iters=12, dtype="f32", k=10,
               sweep=[20, 40, 80, 160], control=10_000, learned=False),
    "c1": dict(desc="10k base x 1k queries, d=128, 20-cluster integer GMM; NN-Descent K=50 L=70 "
                    "S=10 R=100 iters=10; greedy search k=10 from random entries",
               gen="gmm_c1", gen_kw={}, K=50, L=70, S=10, R=100, iters=10, dtype="f32", k=10,
               sweep=[20, 40, 80, 100], control=10_000, learned=False),
}
# NN-Descent leg of the metric: build seconds at graph recall >= 0.95.  K, S
# and the rounds are the workload's; L, R and the candidate slots are the
# fastest of a sweep that still reaches 0.95 (BF16 graph recall 0.953).
NND_LEG = dict(gen="gmm_c2", K=64, L=64, S=16, R=128, iters=12, cand_slots=128, control=10_000,
               desc="1M x 128 Zipf(1.1) GMM, K=64 L=64 S=16 R=128 iters=12, 128 candidate slots per node")
# The kNN graph the refine workload starts from (kept as first measured).
REFINE_KNN = dict(gen="gmm_c2", K=64, L=96, S=16, R=96, iters=12, cand_slots=256, control=10_000,
                  desc="1M x 128 Zipf(1.1) GMM, K=64 L=96 S=16 R=96 iters=12, 256 candidate slots per node")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def recall_at_k(res, gt, k):
    """src/builder.cpp:465-480: 1 - misses / (Q*k), GT truncated to k."""
    hit = (res[:, :k, None] == gt[:, None, :k]).any(axis=2)
    return float(hit.mean())


def graph_recall(ids, deg, gt, control=10):
    """src/builder.cpp:760-776 over the sampled rows (self excluded in gt)."""
    kk = min(control, gt.shape[1])
    m = ids.shape[1]
    valid = np.arange(m)[None, :] < deg[:, None]
    hits = np.zeros(len(gt))
    for j in range(kk):
        hits += ((ids == gt[:, j:j + 1]) & valid).any(axis=1)
    den = np.minimum(np.minimum(kk, deg), gt.shape[1]).astype(np.float64)
    return float(np.mean(np.where(den > 0, hits / np.maximum(den, 1), 0.0)))


MFMA_PEAK_TFLOPS = {"f32": 157.3, "bf16": 2500.0, "fp8": 5000.0}


def join_roofline(st, N, d, kdim, L, dt, traffic_per_launch=None, misses_per_launch=None, line_ceiling=None):
    """Roofline of the local join over the build's joins.  Algorithmic work
    per (node, round): flops 2 d P with P the pairs evaluated; bytes
    (|new| + |old|) (kdim b + 4) for the listed rows (b = 4 / 2 / 1 bytes an
    element) plus 16 L for the pool's keys and flags; the kernel counts P and
    the listed rows.  Attainable = min(MFMA peak, AI x 8 TB/s)."""
    b = {"f32": 4, "bf16": 2, "fp8": 1}[dt]
    rounds = max(int(st["iters_run"]) - 1, 1)
    t = st["join_ms"] / 1e3
    flops = 2.0 * d * st["pairs"]
    alg_bytes = st["rows_gathered"] * (kdim * b + 4) + 16.0 * L * N * rounds
    ai = flops / alg_bytes
    peak = MFMA_PEAK_TFLOPS[dt]
    attain = min(peak, ai * HBM_PEAK_GBS / 1e3)
    tf = flops / t / 1e12
    out = {"bound": "hbm" if attain < peak else "mfma", "kernel": "nnd_join_kernel", "achieved": round(tf, 2),
           "peak": round(attain, 1), "unit": "TFLOP/s", "frac": round(tf / attain, 4),
           "mfma_peak_tflops": peak, "arith_intensity_flop_per_byte": round(ai, 2),
           "alg_bytes_per_launch": int(alg_bytes / rounds), "alg_GBps": round(alg_bytes / t / 1e9, 1),
           "launches": rounds, "kernel_ms_per_launch": round(st["join_ms"] / rounds, 3),
           "rows_gathered": int(st["rows_gathered"]), "pairs": int(st["pairs"])}
    if traffic_per_launch:
        out["traffic"] = int(traffic_per_launch)
        out["traffic_over_alg_bytes"] = round(traffic_per_launch / (alg_bytes / rounds), 3)
    else:
        out["traffic"] = None
    if misses_per_launch and line_ceiling:
        # the join's records, probes, slot updates and row pieces are random
        # lines: their rate against what the GPU sustains for random lines
        got = misses_per_launch / (t / rounds) / 1e9
        out["line_rate"] = {"unit": "G lines/s", "l2_miss_lines_per_launch": int(misses_per_launch),
                            "achieved": round(got, 1), "ceiling": round(line_ceiling, 1),
                            "frac": round(got / line_ceiling, 3),
                            "note": "L2 misses of the join (profiles/nndescent_join_traffic.json) over its "
                                    "live kernel time, against random 16-byte loads from a 1 GiB array "
                                    "(wv_line_rate, measured in this run)"}
    return out


def lds_residency(L, tuning):
    """Wavefronts a CU holds for a plain search launch of pool L as its LDS
    allows (the library's plan: the pool of L rounded up to 64 keys, a
    64-key block and 64 ids, the visited table and the merge buffer, in the
    160 KiB of a CU), at most 16; the kernel's VGPRs can lower it further
    for small pools."""
    P = max(64, (L + 63) // 64 * 64)
    H = tuning.get("visited_slots") or (min(2048, max(1024, 1 << (8 * L - 1).bit_length())) if L <= 256
                                        else (1024 if L <= 512 else 256))
    mb = tuning.get("merge_buffer", 0)
    B = (256 if L >= 2560 else 0) if mb == 0 else (0 if mb == 1 else mb)
    lds = (P * 8 + 64 * 8 + 64 * 4 + H * 4 + B * 8 + 15) // 16 * 16
    return max(1, min(16, (160 * 1024) // lds))


def c2_search(ctx, ix, base, q, g, dg):
    """The 1M x 128 configuration's search: fixed L in {20, 40, 80, 160},
    k = 10, seeded random entries, on the FP32 NN-Descent graph (K=64 S=16
    12 rounds); recall@10 on all 10,000 queries against FP64 exact kNN, ids
    and distances compared with the oracle on the first 500."""
    import oracle  # the checker
    gt, _, _ = ctx.exact_knn(base, q, 10)
    rows = []
    for L in (20, 40, 80, 160):
        ids, dist, ev, hp, st = ix.search(q, L, 10, seed=7)
        rid, rdist, _, _ = oracle.search_greedy(g, dg, base, q[:500], L, 10, seed=7)
        rows.append(dict(L=L, recall_at_10=round(recall_at_k(ids, gt, 10), 4),
                         qps=round(len(q) / (st["kernel_ms"] / 1e3)), dist_evals_per_query=round(float(ev.mean()), 1),
                         ids_bit_exact_vs_oracle_500=bool(np.array_equal(ids[:500], rid) and
                                                          np.array_equal(dist[:500], rdist))))
    log(f"c2 search: {rows}")
    return {"workload": "1M x 128 Zipf GMM, 10,000 queries, FP32 NN-Descent graph K=64 (L=64 S=16 R=128 "
                        "iters=12), fixed-L greedy search k=10 from seeded random entries", "sweep": rows}


def nndescent_leg(ctx, Index, datagen, extras=True):
    """FP32 and BF16 NN-Descent on the 1M x 128 mixture with the leg's
    settings, with the join's roofline; each dtype's build at the fewest
    rounds (from 8) that reach graph recall 0.95; and the BF16 build's graph
    recall at larger candidate tables (towards the reference's exact
    insertion).  Without `extras`, the two 12-round builds and the search
    sweep only (profiling runs)."""
    w = NND_LEG
    base, cq = getattr(datagen, w["gen"])(n_query=10_000)
    N, d = base.shape
    ix = Index(ctx, base)
    ctl = min(w["control"], len(base))
    gt, _, _ = ctx.exact_knn(base, base[:ctl], 10, self_offset=0)
    traffic, misses = {}, {}
    tp = os.path.join(ROOT, "profiles", "nndescent_join_traffic.json")
    if os.path.exists(tp):
        try:
            tj = json.load(open(tp))
            traffic, misses = tj.get("hbm_bytes_per_launch", {}), tj.get("l2_miss_lines_per_launch", {})
        except (ValueError, OSError):
            traffic, misses = {}, {}
    line_ceiling, _ = ctx.line_rate(1 << 30, iters=64)
    log(f"random-line ceiling: {line_ceiling:.1f} G loads/s (1 GiB array)")
    out = {"workload": w["desc"]}
    kdim = {"f32": (d + 3) // 4 * 4, "bf16": (d + 31) // 32 * 32}
    for dt in ("bf16", "f32"):
        t1 = time.time()
        st = ix.build_nndescent(w["K"], w["L"], w["S"], w["R"], w["iters"], seed=1, dtype=dt,
                                cand_slots=w["cand_slots"])
        wall = time.time() - t1
        g, _, dg = ix.get_graph()
        gr = graph_recall(g[:ctl], dg[:ctl], gt)
        out[dt] = dict(build_s=round(st["total_ms"] / 1e3, 4), wall_s=round(wall, 3), graph_recall=round(gr, 4),
                       reached_0_95=gr >= 0.95, init_ms=round(st["init_ms"], 1), sample_ms=round(st["sample_ms"], 1),
                       join_ms=round(st["join_ms"], 1), merge_ms=round(st["merge_ms"], 1),
                       finalize_ms=round(st["finalize_ms"], 1), pairs=int(st["pairs"]),
                       roofline=join_roofline(st, N, d, kdim[dt], w["L"], dt, traffic.get(dt), misses.get(dt),
                                              line_ceiling))
        log(f"nndescent leg {dt}: {st['total_ms']:.1f} ms, graph recall {gr:.4f}, join {out[dt]['roofline']}")
        if dt == "f32":
            out["c2_search"] = c2_search(ctx, ix, base, cq, g, dg)
    out["bf16_speedup_vs_f32"] = round(out["f32"]["build_s"] / out["bf16"]["build_s"], 3)
    out["bf16_join_speedup_vs_f32"] = round(out["f32"]["join_ms"] / out["bf16"]["join_ms"], 3)
    # build seconds at graph recall 0.95: each dtype's fewest rounds reaching it
    at95 = {}
    for dt in ("bf16", "f32") if extras else ():
        for it in range(8, w["iters"] + 1):
            st = ix.build_nndescent(w["K"], w["L"], w["S"], w["R"], it, seed=1, dtype=dt,
                                    cand_slots=w["cand_slots"])
            g, _, dg = ix.get_graph()
            gr = graph_recall(g[:ctl], dg[:ctl], gt)
            if gr >= 0.95 or it == w["iters"]:
                at95[dt] = dict(iters=it, build_s=round(st["total_ms"] / 1e3, 4), join_ms=round(st["join_ms"], 1),
                                graph_recall=round(gr, 4), reached_0_95=gr >= 0.95)
                break
    out["at_graph_recall_0_95"] = at95 or None
    out["bf16_speedup_vs_f32_at_0_95"] = round(at95["f32"]["build_s"] / at95["bf16"]["build_s"], 3) if at95 else None
    log(f"nndescent leg at graph recall 0.95: {at95}")
    slots = {}
    for c in (256, 1024) if extras else ():
        st = ix.build_nndescent(w["K"], w["L"], w["S"], w["R"], w["iters"], seed=1, dtype="bf16", cand_slots=c)
        g, _, dg = ix.get_graph()
        slots[str(c)] = dict(build_s=round(st["total_ms"] / 1e3, 4), graph_recall=round(graph_recall(g[:ctl], dg[:ctl], gt), 4))
    out["bf16_candidate_slots"] = slots
    log(f"nndescent leg slot tables: {slots}")
    ix.close()
    return out


def join_distance_error(base, ids, deg, dtype, sample, seed=0):
    """Max / mean relative error of the join's distance formula on quantised
    rows against FP64, over the graph edges of `sample` random nodes."""
    from weavess_amd.precision import join_formula_rel_err
    nodes = np.random.default_rng(seed).choice(len(base), sample, replace=False)
    return join_formula_rel_err(base, ids, deg, dtype, nodes)


# 128 candidate slots: the automatic 256 (4 L) gave the same graph recall at
# 12 rounds (0.9996) in 2-4 % more time; at 8 rounds 0.962 against 0.967.
C3_SETTINGS = dict(K=64, L=64, S=16, R=64, iters=12, rerank_mult=2, cand_slots=128)


def c3_leg(ctx, Index, datagen, at95_sweep=True):
    """Precision: NN-Descent on 1M x 960 with FP32, BF16 and FP8 joins (FP32
    rerank of each pool's top 2K), graph recall and distance error vs FP64,
    at the workload's 12 rounds and (with at95_sweep) at the fewest rounds
    that reach graph recall 0.95 (each dtype's own count)."""
    t0 = time.time()
    base, _ = datagen.gmm_c3(n_query=1)
    N, d = base.shape
    log(f"c3 data {N}x{d} in {time.time() - t0:.1f}s")
    ix = Index(ctx, base)
    ctl = 10_000
    gt, _, gt_ms = ctx.exact_knn(base, base[:ctl], 10, self_offset=0)
    log(f"c3 exact FP64 10-NN of {ctl} control nodes: {gt_ms:.1f} ms")
    c = C3_SETTINGS
    per = {}
    traffic = {}
    tp = os.path.join(ROOT, "profiles", "nndescent_c3_join_traffic.json")
    if os.path.exists(tp):
        try:
            traffic = json.load(open(tp)).get("hbm_bytes_per_launch", {})
        except (ValueError, OSError):
            traffic = {}
    for dt in ("f32", "bf16", "fp8"):
        st = ix.build_nndescent(c["K"], c["L"], c["S"], c["R"], c["iters"], seed=1, dtype=dt,
                                rerank_mult=c["rerank_mult"], cand_slots=c["cand_slots"])
        ids, dist, deg = ix.get_graph()
        gr = graph_recall(ids[:ctl], deg[:ctl], gt)
        jmax, jmean, npairs = join_distance_error(base, ids, deg, dt, 2000)
        # stored (reranked FP32) graph distances vs FP64
        sel = np.random.default_rng(1).choice(N, 2000, replace=False)
        a_idx = np.repeat(sel, [int(deg[n]) for n in sel])
        b_idx = np.concatenate([ids[n, :deg[n]] for n in sel]).astype(np.int64)
        ex = ((base[a_idx].astype(np.float64) - base[b_idx].astype(np.float64)) ** 2).sum(1)
        stored = np.concatenate([dist[n, :deg[n]] for n in sel]).astype(np.float64)
        ok = ex > 0
        srel = np.abs(stored[ok] - ex[ok]) / ex[ok]
        kd = {"f32": (d + 3) // 4 * 4, "bf16": (d + 31) // 32 * 32, "fp8": (d + 31) // 32 * 32}[dt]
        per[dt] = dict(build_s=round(st["total_ms"] / 1e3, 4), join_ms=round(st["join_ms"], 1),
                       graph_recall=round(gr, 4), roofline=join_roofline(st, N, d, kd, c["L"], dt, traffic.get(dt)),
                       join_distance_rel_err_max=jmax, join_distance_rel_err_mean=jmean,
                       graph_distance_rel_err_max=float(srel.max()),
                       graph_distance_rel_err_mean=float(srel.mean()), error_pairs=npairs)
        log(f"c3 {dt}: {st['total_ms']:.1f} ms, graph recall {gr:.4f}, join distance rel err max {jmax:.3e} "
            f"mean {jmean:.3e}; graph distances (FP32 rerank) max {srel.max():.3e}")
    # build seconds at graph recall 0.95: each dtype's fewest rounds reaching it
    at95 = {}
    for dt in ("f32", "bf16", "fp8") if at95_sweep else ():
        for it in range(2, c["iters"] + 1):
            st = ix.build_nndescent(c["K"], c["L"], c["S"], c["R"], it, seed=1, dtype=dt,
                                    rerank_mult=c["rerank_mult"], cand_slots=c["cand_slots"])
            ids, _, deg = ix.get_graph()
            gr = graph_recall(ids[:ctl], deg[:ctl], gt)
            if gr >= 0.95 or it == c["iters"]:
                at95[dt] = dict(iters=it, build_s=round(st["total_ms"] / 1e3, 4), join_ms=round(st["join_ms"], 1),
                                graph_recall=round(gr, 4), reached_0_95=gr >= 0.95)
                break
    log(f"c3 at graph recall 0.95: {at95}")
    ix.close()
    del base
    return {"workload": f"c3: {N} x {d} non-negative 500-cluster GMM; NN-Descent K={c['K']} L={c['L']} "
                        f"S={c['S']} R={c['R']} iters={c['iters']}, {c['cand_slots']} candidate slots; "
                        f"FP32 rerank of each pool's top 2K",
            "control_nodes": ctl,
            "error_sample": "graph edges of 2,000 random nodes; join formula emulated on the host "
                            "(BF16 RNE; FP8 E4M3 with per-row absmax/448 scale)",
            "per_dtype": per,
            "bf16_speedup_vs_f32": round(per["f32"]["build_s"] / per["bf16"]["build_s"], 3),
            "fp8_speedup_vs_f32": round(per["f32"]["build_s"] / per["fp8"]["build_s"], 3),
            "at_graph_recall_0_95": at95 or None,
            "bf16_speedup_vs_f32_at_0_95": round(at95["f32"]["build_s"] / at95["bf16"]["build_s"], 3)
            if at95 else None}


def run_c3(args, ctx, Index, datagen):
    """The precision configuration alone, as its own JSON line."""
    leg = c3_leg(ctx, Index, datagen, at95_sweep=not args.builds_only)
    out = {"metric": "NN-Descent precision: graph recall and distance error vs FP64 per join dtype",
           "value": leg["per_dtype"]["bf16"]["graph_recall"], "unit": "graph recall@10 (BF16 join)", "n_gpus": 1,
           "higher_is_better": True, "data": "synthetic, seeded (gmm_c3), 1000000x960",
           "config": {"workload": leg["workload"], "control_nodes": leg["control_nodes"],
                      "error_sample": leg["error_sample"]},
           **{k: v for k, v in leg.items() if k not in ("workload", "control_nodes", "error_sample")}}
    print(json.dumps(out), flush=True)


def run_graphs(args, ctx, Index, datagen):
    """How much of the 2M x 96 workload's pool size is the graph's doing:
    the L that recall@10 0.95 and 0.99 need (bisection on 20,000 queries) on
    the bench's NN-Descent graph, a denser NN-Descent build, and the exact
    K=48 kNN graph (FP64 brute force on the GPU, about 100 s)."""
    base, allq = datagen.gmm_c4(n_query=20_000)
    q = allq
    ix = Index(ctx, base)
    gt, _, _ = ctx.exact_knn(base, q, 10)
    ctl = 10_000
    gtg, _, _ = ctx.exact_knn(base, base[:ctl], 10, self_offset=0)

    def l_at(target):
        cache = {}

        def rec(L):
            if L not in cache:
                ids, _, ev, hp, st = ix.search(q, L, 10, seed=7)
                cache[L] = (recall_at_k(ids, gt, 10), float(ev.mean()), float(hp.mean()), st["kernel_ms"])
            return cache[L][0]
        lo, hi = 64, L_MAX
        if rec(hi) < target:
            return None
        while hi - lo > max(1, hi // 64):
            mid = (lo + hi) // 2
            if rec(mid) >= target:
                hi = mid
            else:
                lo = mid
        r, ev, hp, ms = cache[hi]
        return dict(L=hi, recall=round(r, 4), evals=round(ev), hops=round(hp), qps=round(len(q) / (ms / 1e3)))

    res = []

    def study(name, extra):
        g, _, dg = ix.get_graph()
        row = dict(graph=name, graph_recall=round(graph_recall(g[:ctl], dg[:ctl], gtg), 4), **extra)
        for tgt in (0.95, 0.99):
            row[f"at_{tgt:g}"] = l_at(tgt)
        log(f"graphs: {row}")
        res.append(row)

    for (L, S, R, it) in ((128, 32, 128, 20), (256, 48, 256, 24)):
        st = ix.build_nndescent(48, L, S, R, it, seed=1, dtype="bf16")
        study(f"NN-Descent K=48 L={L} S={S} R={R} iters={it} BF16", dict(build_s=round(st["total_ms"] / 1e3, 3)))
    t1 = time.time()
    K = 48
    ids = np.zeros((len(base), K), np.uint32)
    step = 250_000
    for s0 in range(0, len(base), step):
        ids[s0:s0 + step], _, _ = ctx.exact_knn(base, base[s0:s0 + step], K, self_offset=s0)
        log(f"exact K=48 graph: {min(s0 + step, len(base))} rows, {time.time() - t1:.1f} s")
    ix.set_graph(ids, np.full(len(base), K, np.uint32))
    study("exact K=48 kNN graph (FP64)", dict(build_s=round(time.time() - t1, 1)))
    ix.close()
    out = {"metric": "pool size L needed for recall@10 0.95 / 0.99 per graph", "value": res[-1]["at_0.95"]["L"]
           if res[-1]["at_0.95"] else None, "unit": "L (exact K=48 graph, recall 0.95)", "n_gpus": 1,
           "higher_is_better": False, "data": "synthetic, seeded (gmm_c4), 2000000x96, 20000 queries",
           "config": {"workload": "graphs: 2M x 96 unit-norm GMM; greedy search k=10 from random entries"},
           "graphs": res}
    print(json.dumps(out), flush=True)


def run_nnd(args, ctx, Index, datagen):
    """The NN-Descent leg alone (FP32 and BF16 builds of the 1M x 128 mixture)."""
    leg = nndescent_leg(ctx, Index, datagen, extras=not args.builds_only)
    out = {"metric": "NN-Descent build s @ 0.95 graph recall", "value": leg["bf16"]["build_s"], "unit": "s (BF16 join)",
           "n_gpus": 1, "higher_is_better": False, "dtype": "bf16", "data": "synthetic, seeded (gmm_c2)",
           "config": {"workload": "nnd: " + NND_LEG["desc"]}, "nndescent": leg}
    print(json.dumps(out), flush=True)


def run_c4(args, ctx, Index, datagen):
    """Batching: one launch over 200,000 queries with per-query L ~ U{16..400}
    and k in {1, 10, 100} (L_eff = max(L, k)) on the 2M x 96 graph."""
    from weavess_amd.devmem import DeviceArray, pad_rows, synchronize
    t0 = time.time()
    base, queries = datagen.gmm_c4(n_query=200_000)
    N, d = base.shape
    Q = len(queries)
    Lq, kq = datagen.c4_query_budgets(Q)
    log(f"data {N}x{d}, {Q} queries in {time.time() - t0:.1f}s")
    ix = Index(ctx, base)
    st = ix.build_nndescent(48, 128, 32, 128, 20, seed=1, dtype="bf16")
    log(f"graph: {st['total_ms']:.1f} ms")
    ns = 20_000
    gt, _, _ = ctx.exact_knn(base, queries[:ns], 100)
    kmax = 100
    dq = DeviceArray.from_host(pad_rows(queries))
    dL = DeviceArray.from_host(Lq)
    dk = DeviceArray.from_host(kq)
    d_ids = DeviceArray((Q, kmax), np.uint32)
    d_ev = DeviceArray((Q,), np.uint32)
    d_hp = DeviceArray((Q,), np.uint32)

    def run():
        return ix.search_dev(dq.ptr, Q, 0, 0, d_ids.ptr, seed=7, d_L=dL.ptr, d_k=dk.ptr, k_stride=kmax,
                             d_evals=d_ev.ptr, d_hops=d_hp.ptr)

    tuning, _ = ix.autotune_dev(dq.ptr, 8192, int(Lq.max()), 100, seed=7)
    log(f"autotuned at L={int(Lq.max())}, k=100: {tuning}")
    for _ in range(args.warmup):
        run()
    synchronize()
    kern = []
    t1 = time.perf_counter()
    for _ in range(args.steps):
        s_ = run()
        kern.append(s_["kernel_ms"])
    synchronize()
    ms = (time.perf_counter() - t1) * 1e3 / args.steps
    ids = d_ids.to_host()[:ns]
    rec = {}
    for kk in (1, 10, 100):
        m = kq[:ns] == kk
        hit = (ids[m][:, :kk, None] == gt[m][:, None, :kk]).any(axis=2)
        rec[f"k={kk}"] = dict(queries=int(m.sum()), recall=round(float(hit.mean()), 4))
    ev, hp = d_ev.to_host(), d_hp.to_host()
    out = {"metric": "QPS, batched search with per-query L and k", "value": round(Q / (ms / 1e3), 1),
           "unit": "QPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
           "higher_is_better": True, "dtype": "f32", "data": f"synthetic, seeded (gmm_c4), {N}x{d}, {Q} queries",
           "config": {"workload": "c4: 2M x 96 unit-norm GMM; NN-Descent K=48 (L=128 S=32 R=128 iters=20, BF16); "
                                  "per-query L ~ U{16..400}, k in {1,10,100}, L_eff = max(L, k); one launch",
                      "tuning": tuning},
           "kernel_ms": round(float(np.mean(kern)), 3), "recall_by_k": rec,
           "dist_evals_per_query": round(float(ev.mean()), 1), "hops_per_query": round(float(hp.mean()), 1)}
    print(json.dumps(out), flush=True)
    for a in (dq, dL, dk, d_ids, d_ev, d_hp):
        a.free()
    ix.close()


def run_refine(args, ctx, Index, datagen):
    """Refiners downstream of NN-Descent on the 1M x 128 mixture of the
    NN-Descent leg: NSG, SSG, DPG and Vamana (two passes from a random
    graph), each searched from its own entries; EFANNA's NN-Descent against
    KGraph's; the CPU restatement of NSG timed on a 20k-node sample."""
    from weavess_amd.entries import centroid_ep, centroid_entries, subcentroid_entries
    w = REFINE_KNN
    base, q = getattr(datagen, w["gen"])(n_query=10_000)
    N, d = base.shape
    ix = Index(ctx, base)
    gt, _, _ = ctx.exact_knn(base, q, 10)
    ctl = 10_000
    gtg, _, _ = ctx.exact_knn(base, base[:ctl], 10, self_offset=0)
    kst = ix.build_nndescent(w["K"], w["L"], w["S"], w["R"], w["iters"], seed=1, dtype="bf16",
                             cand_slots=w["cand_slots"])
    kid, _, kdeg = ix.get_graph()
    ep = centroid_ep(ctx, base)
    log(f"refine: kNN graph {kst['total_ms']:.1f} ms, ep {ep}")
    row_bytes = ((d + 3) // 4) * 16

    def sweep(g, dg, entries, Ls):
        ix.set_graph(g, dg)
        out = []
        for L in Ls:
            if entries == "rand":
                ids, _, ev, _, ss = ix.search(q, L, 10, seed=0)
            else:
                ent = entries(g, dg, L)
                ids, _, ev, _, ss = ix.search(q, L, 10, entries=ent)
            out.append(dict(L=L, recall=round(recall_at_k(ids, gt, 10), 4),
                            qps=round(len(q) / (ss["kernel_ms"] / 1e3)), dist_evals=round(float(ev.mean()))))
        return out

    def rounded(st):
        return {k: (round(v, 1) if isinstance(v, float) else int(v)) for k, v in st.items()}

    cent = lambda g, dg, L: centroid_entries(ep, g, dg, len(q), L, N)
    res = {"knn_graph": dict(build_s=round(kst["total_ms"] / 1e3, 4), search=sweep(kid, kdeg, "rand", [80, 160]))}
    # NSG (L_refine 150, R 30, C 400: the reference's sift1M setting)
    ix.set_graph(kid, kdeg)
    st = ix.build_nsg(150, 30, 400, ep, seed=1)
    g, _, dg = ix.get_graph()
    st = rounded(st)
    st["candidate_search_gather_GBps"] = round(st["evals"] * row_bytes / (st["search_ms"] / 1e3) / 1e9, 1)
    res["nsg"] = dict(params="L_refine=150 R=30 C=400", stats=st, deg_mean=round(float(dg.mean()), 2),
                      search=sweep(g, dg, cent, [40, 80, 160]))
    log(f"refine nsg {st}")
    # SSG (L_refine 100, R 32, A 60, n_try 10)
    ix.set_graph(kid, kdeg)
    st, eps = ix.build_ssg(100, 32, 60.0, 10, seed=1)
    g, _, dg = ix.get_graph()
    res["ssg"] = dict(params="L_refine=100 R=32 A=60 n_try=10", stats=rounded(st), deg_mean=round(float(dg.mean()), 2),
                      search=sweep(g, dg, lambda g_, d_, L: subcentroid_entries(eps, len(q), L, N), [40, 80, 160]))
    log(f"refine ssg {res['ssg']['stats']}")
    # DPG (K / 2 = 32 a node, mirrored)
    ix.set_graph(kid, kdeg)
    st = ix.build_dpg(w["K"] // 2)
    g, _, dg = ix.get_graph()
    res["dpg"] = dict(params=f"range={w['K'] // 2}", stats=rounded(st), deg_mean=round(float(dg.mean()), 2),
                      search=sweep(g, dg, "rand", [40, 80, 160]))
    log(f"refine dpg {res['dpg']['stats']}")
    # Vamana: INIT_RAND (32 random neighbours), passes alpha 1 and 2 (L_refine 70, R 32)
    t0 = time.time()
    ix.build_nndescent(32, 32, 32, 0, 0, seed=1)
    sts = [rounded(ix.build_vamana(70, 32, 1024, ep, a, seed=1)) for a in (1.0, 2.0)]
    g, _, dg = ix.get_graph()
    res["vamana"] = dict(params="INIT_RAND 32, L_refine=70 R=32 C=1024, alpha 1 then 2", passes=sts,
                         wall_s=round(time.time() - t0, 3), deg_mean=round(float(dg.mean()), 2),
                         search=sweep(g, dg, cent, [40, 80, 160]))
    log(f"refine vamana {sts}")
    # EFANNA vs KGraph NN-Descent (BF16, same K, L, S, R, slots)
    ef = {}
    for name, fn in (("kgraph", ix.build_nndescent), ("efanna", ix.build_efanna)):
        for it in (10, 11, 12):
            st = fn(w["K"], w["L"], w["S"], w["R"], it, seed=1, dtype="bf16", cand_slots=w["cand_slots"])
            gg, _, gd = ix.get_graph()
            ef[f"{name}_{it}"] = dict(build_s=round(st["total_ms"] / 1e3, 4),
                                      graph_recall=round(graph_recall(gg[:ctl], gd[:ctl], gtg), 4))
    res["efanna_vs_kgraph"] = ef
    log(f"refine efanna {ef}")
    ix.close()
    cpu = None
    if not args.no_cpu:
        import oracle as O
        from weavess_amd.devmem import host_cores
        sub = base[:20_000]
        ix2 = Index(ctx, sub)
        ix2.build_nndescent(w["K"], w["L"], w["S"], w["R"], w["iters"], seed=1, dtype="bf16",
                            cand_slots=w["cand_slots"])
        k2, _, d2 = ix2.get_graph()
        ep2 = centroid_ep(ctx, sub)
        st2 = ix2.build_nsg(150, 30, 400, ep2, seed=1)
        g2, _, gd2 = ix2.get_graph()
        ix2.close()
        t1 = time.time()
        oi, _, odeg, _ = O.nsg(sub, k2, d2, 150, 30, 400, ep2, seed=1)
        cpu_s = time.time() - t1
        same = bool(np.array_equal(gd2, odeg) and all(g2[i, :gd2[i]].tolist() == oi[i, :odeg[i]].tolist()
                                                      for i in range(len(sub))))
        cpu = dict(value=round(cpu_s, 3), unit="s (NSG refine of 20,000 nodes)", cores=host_cores(), kind="port",
                   sample="first 20,000 points, their BF16 NN-Descent graph, L_refine=150 R=30 C=400",
                   gpu_s=round(st2["total_ms"] / 1e3, 4), graph_bit_exact_vs_gpu=same)
        log(f"refine cpu baseline {cpu}")
    out = {"metric": "refiners of the NN-Descent graph: build seconds and search recall/QPS per graph",
           "value": res["nsg"]["stats"]["total_ms"] / 1e3, "unit": "s (NSG refine, 1M x 128)", "n_gpus": 1,
           "higher_is_better": False, "data": f"synthetic, seeded ({w['gen']}), {N}x{d}, 10000 queries",
           "config": {"workload": "refine: 1M x 128 Zipf(1.1) GMM; BF16 NN-Descent K=64 L=96 S=16 R=96 iters=12; "
                                  "refiners on the GPU; greedy search k=10"},
           "results": res, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def batched_leg(ctx, ix, base, queries, dq, datagen, args):
    """The batching configuration on the 2M x 96 graph: the 200,000 queries
    with per-query L ~ U{16..400} and k in {1, 10, 100} in one launch, as the
    configuration states it: captured once in a hipGraph and replayed, with
    the autotuner's plan (schedule, gather tile, visited table, buffer,
    wavefronts per workgroup and per query, tuned under the queries' own
    budgets) and, beside it, with the persistent work-queue schedule
    (largest L first).  Every replay's ids are
    checked against the eager launch's; recall@k per k on 4,000 queries."""
    from weavess_amd.devmem import DeviceArray, synchronize
    Q = len(queries)
    Lq, kq = datagen.c4_query_budgets(Q)
    dL, dk = DeviceArray.from_host(Lq), DeviceArray.from_host(kq)
    d_ids = DeviceArray((Q, 100), np.uint32)
    t1 = time.time()
    # tuned under the workload's own budgets, on the first 65,536 queries
    # (with 8,192 the short launches left the pick between the 4- 8- and
    # 16-group tiles to noise: 1.42 against 1.75 M QPS on the full set)
    tuning, _ = ix.autotune_budgets_dev(dq.ptr, min(Q, 65536), dL.ptr, dk.ptr, 100, seed=7)
    log(f"batched: autotuned {tuning} ({time.time() - t1:.1f} s)")

    def eager():
        return ix.search_dev(dq.ptr, Q, 0, 0, d_ids.ptr, seed=7, d_L=dL.ptr, d_k=dk.ptr, k_stride=100)
    eager()
    synchronize()
    ref_ids = d_ids.to_host()
    steps = min(args.steps, 3)
    kern = []
    t1 = time.perf_counter()
    for _ in range(steps):
        kern.append(eager()["kernel_ms"])
    synchronize()
    ms = (time.perf_counter() - t1) * 1e3 / steps

    def captured(cfg):
        ix.set_search_config(**cfg)
        g = ix.search_graph(dq.ptr, Q, 0, 0, d_ids.ptr, seed=7, d_L=dL.ptr, d_k=dk.ptr, k_stride=100)
        try:
            g.launch(sync=True)  # warm-up replay
            t0 = time.perf_counter()
            for _ in range(steps):
                g.launch(sync=False)
            synchronize()
            gms = (time.perf_counter() - t0) * 1e3 / steps
        finally:
            g.close()
        same = bool(np.array_equal(d_ids.to_host(), ref_ids))
        return dict(qps=round(Q / (gms / 1e3), 1), ms_per_step=round(gms, 3), ids_equal_eager=same,
                    tuning=dict(cfg))
    graph_auto = captured(tuning)
    pers_cfg = dict(tuning, schedule=2, waves_per_query=0)
    graph_pers = captured(pers_cfg)
    ix.set_search_config(**tuning)
    ns = 4000
    gt, _, _ = ctx.exact_knn(base, queries[:ns], 100)
    ids = ref_ids[:ns]
    rec = {}
    for kk in (1, 10, 100):
        m = kq[:ns] == kk
        rec[f"k={kk}"] = dict(queries=int(m.sum()), recall=round(recall_at_k(ids[m], gt[m], kk), 4))
    for a_ in (dL, dk, d_ids):
        a_.free()
    best = max((graph_auto, graph_pers), key=lambda x: x["qps"])
    out = dict(workload="2M x 96, 200,000 queries, per-query L ~ U{16..400}, k in {1,10,100}, one launch",
               qps=best["qps"], ms_per_step=best["ms_per_step"], served="hipGraph replay, " +
               ("persistent work queue" if best is graph_pers else "autotuned plan"),
               graph_autotuned=graph_auto, graph_persistent=graph_pers,
               eager=dict(qps=round(Q / (ms / 1e3), 1), ms_per_step=round(ms, 3),
                          kernel_ms=round(float(np.mean(kern)), 3)),
               steps=steps, tuning=tuning, recall_by_k=rec)
    log(f"batched: {out}")
    return out


def learned_leg(ctx, ix, allq, n_work, w, L_plain, target, k, run_learned, timed, steps, warmup, gt, ns, Q,
                n_tune, dq, base):
    """The learned stop at one recall target, beside plain fixed L = L_plain.

    Traces of plain search at the pool cap 1.3 x L_plain on the training
    queries (at most 500,000 per-expansion rows) train the 8 -> 32 -> 1
    network on the GPU; the threshold is calibrated on the calibration
    queries (target below); the search is timed at that cap.  Beside it,
    the ceiling of any stop rule at cap L_plain: a rule that stops every
    query exactly at its settle hop (the last expansion that changed its
    top-k) returns plain's results after that many expansions.  The
    calibration target is the recall target plus 0.003, or plus a tenth of
    1 - target where that is smaller."""
    from weavess_amd import learned
    tq = allq[n_work:n_work + w["n_train"]]
    cq = allq[n_work + w["n_train"]:n_work + w["n_train"] + w["n_cal"]]
    gt_cal, _, _ = ctx.exact_knn(base, cq, k)
    # ceiling of stop rules at plain's pool
    ids, _, ev, hp, _, _, _, settle = ix.search_ext(cq, L_plain, k, seed=5, want_settle=True)
    ceiling = dict(cap_L=L_plain, queries=len(cq), recall_at_10=round(recall_at_k(ids, gt_cal, k), 4),
                   hops_per_query=round(float(hp.mean()), 1), settle_hop_mean=round(float(settle.mean()), 1),
                   settle_over_hops=round(float(settle.mean() / hp.mean()), 4),
                   settle_fraction_pct_10_25_50_75_90=[round(float(np.percentile(settle / hp, p_)), 3)
                                                      for p_ in (10, 25, 50, 75, 90)],
                   max_speedup_of_any_stop_rule=round(float(hp.mean() / max(settle.mean(), 1.0)), 3))
    log(f"stop-rule ceiling at L={L_plain}: {ceiling}")
    cap = min(L_MAX, int(round(1.3 * L_plain)))
    t1 = time.time()
    X, y, _ = learned.collect_traces(ix, tq, cap, k, seed=3)
    if len(X) > 500_000:
        sel = np.sort(np.random.default_rng(0).choice(len(X), 500_000, replace=False))
        X, y = X[sel], y[sel]
    model, tst = learned.train(ctx, X, y, epochs=20, batch=4096, seed=1)
    # calibration margin: 0.003, or a tenth of the misses the target allows
    # when that is smaller (at 0.99 the cap's own recall is about 0.991)
    cal_target = target + min(0.003, 0.1 * (1.0 - target))
    cal = learned.calibrate(ix, cq, gt_cal, model, cap, k, cal_target, seed=5,
                            thresholds=np.arange(-4.0, 12.01, 0.25), bisect=True)
    log(f"learned at {target}: {len(X)} traces, training {tst['kernel_ms']:.1f} ms device, loss "
        f"{tst['first_loss']:.4f} -> {tst['last_loss']:.4f}; calibration {cal}; {time.time() - t1:.1f} s")
    info = dict(cap_L=cap, traces=int(len(X)), train_ms=round(tst["kernel_ms"], 2),
                loss=[round(tst["first_loss"], 4), round(tst["last_loss"], 4)],
                threshold=None if cal is None else cal[0],
                calibration_recall=None if cal is None else round(cal[1], 4))
    if cal is not None:
        model.logit_threshold = cal[0]
        t1 = time.time()
        info["tuning"], _ = ix.autotune_dev(dq.ptr, n_tune, cap, k, seed=7, model=model)
        log(f"autotuned learned search at cap {cap}: {info['tuning']} ({time.time() - t1:.1f} s)")
        l_ms, l_kern, l_ev, l_hp = timed(lambda: run_learned(cap, model), steps, warmup)
        l_rec = learned.recall_at_k(run_learned.ids()[:ns], gt, k)
        info.update(qps=round(Q / (l_ms / 1e3), 1), ms_per_step=round(l_ms, 3), kernel_ms=round(l_kern, 3),
                    recall_at_10=round(l_rec, 4), target_reached=l_rec >= target,
                    dist_evals_per_query=round(l_ev / Q, 1), hops_per_query=round(l_hp / Q, 1), steps=steps,
                    _ev=l_ev, _hp=l_hp)
        log(f"learned timed at {target}: {info['qps']:.0f} QPS, recall {l_rec:.4f}")
    return {"learned": info, "stop_ceiling": ceiling}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c5", choices=sorted(WORKLOADS) + ["c3", "c4", "refine", "nnd", "graphs"])
    ap.add_argument("--target", type=float, default=0.95)
    ap.add_argument("--no-cpu", action="store_true", help="skip the host-core baseline")
    ap.add_argument("--no-nnd-leg", action="store_true", help="skip the 1M x 128 NN-Descent leg")
    ap.add_argument("--no-c3-leg", action="store_true", help="skip the 1M x 960 precision leg")
    ap.add_argument("--builds-only", action="store_true",
                    help="--workload c3 / nnd: the 12-round build of each dtype only (profiling runs)")
    ap.add_argument("--fixed-L", type=int, default=0,
                    help="profiling run: plain search at this L only (no L sweep, learned stop, second "
                         "target, NN-Descent legs or batching leg)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(queries would shard over replicas with no collective)", file=sys.stderr)
        sys.exit(2)

    from weavess_amd import Context, Index, datagen, learned
    from weavess_amd.devmem import DeviceArray, host_cores, pad_rows, synchronize

    ctx = Context(0)
    if args.workload in ("c3", "c4", "refine", "nnd", "graphs"):
        {"c3": run_c3, "c4": run_c4, "refine": run_refine, "nnd": run_nnd,
         "graphs": run_graphs}[args.workload](args, ctx, Index, datagen)
        ctx.close()
        return
    w = dict(WORKLOADS[args.workload])
    if args.fixed_L:
        w.update(sweep=[args.fixed_L], learned=False)
    nnd_leg = None
    if not args.no_nnd_leg and args.workload == "c5" and not args.fixed_L:
        nnd_leg = nndescent_leg(ctx, Index, datagen)
    c3 = None
    if not args.no_c3_leg and args.workload == "c5" and not args.fixed_L:
        c3 = c3_leg(ctx, Index, datagen)

    t0 = time.time()
    base, allq = getattr(datagen, w["gen"])(**w["gen_kw"])
    N, d = base.shape
    k = w["k"]
    n_work = w.get("n_work", len(allq))
    queries = allq[:n_work]
    Q = n_work
    log(f"data {N}x{d}, {Q} workload queries in {time.time() - t0:.1f}s")

    ix = Index(ctx, base)
    ctl = min(w["control"], N)
    gt_graph, _, _ = ctx.exact_knn(base, base[:ctl], 10, self_offset=0)
    st = ix.build_nndescent(w["K"], w["L"], w["S"], w["R"], w["iters"], seed=1, dtype=w["dtype"])
    graph, gdist, gdeg = ix.get_graph()
    gr = graph_recall(graph[:ctl], gdeg[:ctl], gt_graph)
    graph_info = dict(K=w["K"], L=w["L"], S=w["S"], R=w["R"], iters=w["iters"], dtype=w["dtype"],
                      build_s=round(st["total_ms"] / 1e3, 4), graph_recall=round(gr, 4))
    log(f"graph: nndescent {w['dtype']} {st['total_ms']:.1f} ms, graph recall {gr:.4f}")

    ns = min(Q, w.get("recall_sample", Q))
    gt, _, gt_ms = ctx.exact_knn(base, queries[:ns], k)
    log(f"exact FP64 ground truth for {ns} queries: {gt_ms:.1f} ms")

    dq = DeviceArray.from_host(pad_rows(queries))
    d_ids = DeviceArray((Q, k), np.uint32)
    d_dist = DeviceArray((Q, k), np.float32)
    d_ev = DeviceArray((Q,), np.uint32)
    d_hp = DeviceArray((Q,), np.uint32)

    def run_plain(L, nq=Q):
        return ix.search_dev(dq.ptr, nq, L, k, d_ids.ptr, d_dist.ptr, seed=7, d_evals=d_ev.ptr, d_hops=d_hp.ptr)

    def run_learned(cap, model, nq=Q):
        return ix.search_ext_dev(dq.ptr, nq, cap, k, d_ids.ptr, model=model, d_out_dist=d_dist.ptr, seed=7,
                                 d_evals=d_ev.ptr, d_hops=d_hp.ptr)
    run_learned.ids = d_ids.to_host

    cache = {}

    def measure(L):
        if L not in cache:
            s = run_plain(L, ns)
            r = recall_at_k(d_ids.to_host()[:ns], gt, k)
            cache[L] = dict(L=L, recall=r, qps=ns / (s["kernel_ms"] / 1e3), evals=s["dist_evals"] / ns)
            log(f"plain L={L}: recall@{k} {r:.4f}, {cache[L]['qps']:.0f} QPS (kernel), "
                f"{cache[L]['evals']:.0f} evals/query")
        return cache[L]

    sweep = [measure(L) for L in w["sweep"]]

    def l_for(target):
        """Smallest L (to 1/64) whose sample recall reaches the target."""
        below = [L for L, c in cache.items() if c["recall"] < target]
        above = [L for L, c in cache.items() if c["recall"] >= target]
        lo = max(below) if below else k
        hi = min(above) if above else None
        if hi is None:
            hi = max(cache)
            while measure(hi)["recall"] < target and hi < L_MAX:
                lo, hi = hi, min(L_MAX, hi * 2)
            if measure(hi)["recall"] < target:
                return hi, False
        while hi - lo > max(1, hi // 64):
            mid = (lo + hi) // 2
            if measure(mid)["recall"] >= target:
                hi = mid
            else:
                lo = mid
        return hi, True

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        synchronize()
        kern, ev, hp = [], 0, 0
        t1 = time.perf_counter()
        for _ in range(steps):
            s = fn()
            kern.append(s["kernel_ms"])
            ev, hp = s["dist_evals"], s["hops"]
        synchronize()
        t2 = time.perf_counter()
        return (t2 - t1) * 1e3 / steps, float(np.mean(kern)), ev, hp

    n_tune = min(Q, 8192)
    targets = [args.target] + ([0.99] if w["learned"] and args.target < 0.99 else [])
    per_target = {}
    primary = None
    for ti, tgt in enumerate(targets):
        main_t = ti == 0
        steps, warm = (args.steps, args.warmup) if main_t else (min(args.steps, 2), min(args.warmup, 1))
        L_t, reached = (args.fixed_L, measure(args.fixed_L)["recall"] >= tgt) if args.fixed_L else l_for(tgt)
        op = measure(L_t)
        log(f"plain operating point for {tgt}: L={L_t}, recall {op['recall']:.4f}" +
            ("" if reached else f" (target not reached at L <= {L_MAX})"))
        t1 = time.time()
        tuning, _ = ix.autotune_dev(dq.ptr, n_tune, L_t, k, seed=7)
        log(f"autotuned plain search at L={L_t}: {tuning} ({time.time() - t1:.1f} s)")
        p_ms, p_kern, p_ev, p_hp = timed(lambda: run_plain(L_t), steps, warm)
        p_rec = recall_at_k(d_ids.to_host()[:ns], gt, k)
        plain_t = dict(L=L_t, qps=round(Q / (p_ms / 1e3), 1), ms_per_step=round(p_ms, 3),
                       kernel_ms=round(p_kern, 3), recall_at_10=round(p_rec, 4), target_reached=reached,
                       dist_evals_per_query=round(p_ev / Q, 1), hops_per_query=round(p_hp / Q, 1),
                       tuning=tuning, steps=steps)
        log(f"plain timed at {tgt}: {plain_t['qps']:.0f} QPS at L={L_t}")
        entry = {"plain": plain_t}
        if main_t:
            primary = dict(L=L_t, reached=reached, tuning=tuning, ms=p_ms, kern=p_kern, ev=p_ev, hp=p_hp,
                           rec=p_rec)
        if w["learned"]:
            entry.update(learned_leg(ctx, ix, allq, n_work, w, L_t, tgt, k, run_learned, timed, steps, warm,
                                     gt, ns, Q, n_tune, dq, base))
            if "qps" in entry["learned"]:
                entry["learned_vs_plain_qps"] = round(entry["learned"]["qps"] / plain_t["qps"], 3)
        per_target[f"{tgt:g}"] = entry
        ix.set_search_config(**primary["tuning"])

    first = per_target[f"{targets[0]:g}"]
    plain = first["plain"]
    L_plain, reached, plain_tuning = primary["L"], primary["reached"], primary["tuning"]
    # the served method: the faster of plain and learned among those that
    # reach the target recall on the sample (plain is the fallback)
    ms_step, kern_ms, evals, hops = primary["ms"], primary["kern"], primary["ev"], primary["hp"]
    rec, method = primary["rec"], "plain fixed-L"
    lrn = first.get("learned")
    if lrn and lrn.get("target_reached") and (lrn["ms_per_step"] < ms_step or not reached):
        ms_step, kern_ms, evals, hops = lrn["ms_per_step"], lrn["kernel_ms"], lrn["_ev"], lrn["_hp"]
        rec, method = lrn["recall_at_10"], "learned"
    for e in per_target.values():
        if "learned" in e:
            e["learned"].pop("_ev", None)
            e["learned"].pop("_hp", None)
    value = Q / (ms_step / 1e3)

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "QPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": f"synthetic, seeded ({w['gen']}), {N}x{d} base, {Q} queries",
        "config": {"workload": args.workload + ": " + w["desc"], "k": k, "method": method,
                   "methods_compared": "plain fixed-L" + (" and learned stop" if w["learned"] else ""),
                   "recall_at_10": round(rec, 4), "target_recall": args.target,
                   "recall_sample": ns, "entries": "seeded GenRandom per query", "graph": graph_info},
        "roofline": None,
        "plain": plain,
        "l_sweep": [dict(L=c["L"], recall=round(c["recall"], 4), qps=round(c["qps"], 1))
                    for c in sorted(cache.values(), key=lambda c: c["L"])],
    }
    if "learned" in first:
        out["learned"] = first["learned"]
        if "learned_vs_plain_qps" in first:
            out["learned_vs_plain_qps"] = first["learned_vs_plain_qps"]
    if len(per_target) > 1 or w["learned"]:
        out["targets"] = per_target
    if nnd_leg is not None:
        out["nndescent"] = nnd_leg
    if c3 is not None:
        out["nndescent_c3"] = c3

    if not args.no_cpu:
        import oracle  # the CPU baseline leg: the oracle is timed here, never the product
        cores = host_cores()
        # bounded sample: about 10-30 s of host work at this L
        probe = min(ns, 64)
        c0 = time.perf_counter()
        oracle.search_greedy(graph, gdeg, base, queries[:probe], L_plain, k, seed=7, nthreads=cores)
        per_q = (time.perf_counter() - c0) / probe
        sample = int(min(ns, max(probe, 15.0 / max(per_q, 1e-9))))
        times = []
        res = None
        while len(times) < 3:
            c0 = time.perf_counter()
            res, _, cev, _ = oracle.search_greedy(graph, gdeg, base, queries[:sample], L_plain, k, seed=7,
                                                  nthreads=cores)
            times.append(time.perf_counter() - c0)
            if sum(times) > 30:
                break
        cpu_qps = sample / float(np.median(times))
        ix.set_search_config(**plain_tuning)  # the timed plain launch's tuning
        run_plain(L_plain)
        gpu_ids = d_ids.to_host()[:sample]
        gpu_ev = d_ev.to_host()[:sample]
        out["cpu_baseline"] = {"value": round(cpu_qps, 1), "unit": "QPS", "cores": cores, "kind": "port",
                               "sample": f"{sample} workload queries, plain fixed L={L_plain}, median of "
                                         f"{len(times)} runs (oracle greedy search, OpenMP over queries, -O3)",
                               "ids_bit_exact_vs_gpu": bool(np.array_equal(res, gpu_ids)),
                               "gpu_evals_over_reference_evals": round(float(gpu_ev.sum()) / float(cev.sum()), 4)}
        log(f"cpu baseline {cpu_qps:.0f} QPS on {cores} threads ({sample} queries); ids equal: "
            f"{out['cpu_baseline']['ids_bit_exact_vs_gpu']}")
        # one thread, as the reference's KGraph search loop runs (src/builder.cpp:445-454)
        s1 = min(sample, 32)
        c0 = time.perf_counter()
        oracle.search_greedy(graph, gdeg, base, queries[:s1], L_plain, k, seed=7, nthreads=1)
        out["cpu_baseline"]["single_thread_qps"] = round(s1 / (time.perf_counter() - c0), 2)
        out["cpu_baseline"]["single_thread_sample"] = f"{s1} workload queries"
        # NN-Descent build on the host cores at the 10k x 128 configuration
        # (InitRandom + NN-Descent + finalize, the reference's "Build cost"),
        # timed with the reference's exact bounded insertion
        # (include/weavess/index.h:120-135, a mutex per node); beside it the
        # reference's single-thread sweep of the reverse test and the slot
        # table the GPU uses, all against the FP64 exact kNN
        w1 = WORKLOADS["c1"]
        b1, _ = getattr(datagen, w1["gen"])(**w1["gen_kw"])
        ctl1 = len(b1)
        gt1, _, _ = ctx.exact_knn(b1, b1, 10, self_offset=0)
        prm = (w1["K"], w1["L"], w1["S"], w1["R"], w1["iters"])
        ix1 = Index(ctx, b1)
        gpu_rec, gpu_s = [], []
        for sd in (1, 2, 3):
            st1 = ix1.build_nndescent(*prm, seed=sd, dtype="f32")
            g1, _, d1 = ix1.get_graph()
            gpu_rec.append(graph_recall(g1, d1, gt1))
            gpu_s.append(st1["total_ms"] / 1e3)
            if sd == 1:
                g1_first = g1
        ix1.close()
        cpu_times, cpu_rec, sweep_rec = [], [], []
        for sd in (1, 2, 3):
            c0 = time.perf_counter()
            o1, _, od1 = oracle.nndescent(b1, *prm, seed=sd, nthreads=cores)
            cpu_times.append(time.perf_counter() - c0)
            cpu_rec.append(graph_recall(o1, od1, gt1))
            s1, _, sd1 = oracle.nndescent(b1, *prm, seed=sd, nthreads=cores, sweep=True)
            sweep_rec.append(graph_recall(s1, sd1, gt1))
        so, _, _ = oracle.nndescent(b1, *prm, seed=1, cand_slots=st1["cand_slots"], nthreads=cores)
        out["cpu_baseline"]["nndescent_10k"] = {
            "workload": f"10k x 128 integer GMM, K={w1['K']} L={w1['L']} S={w1['S']} R={w1['R']} iters={w1['iters']}",
            "cpu_s": round(float(np.median(cpu_times)), 3), "cpu_rule": "exact bounded insertion (index.h:120-135)",
            "gpu_s": round(float(np.median(gpu_s)), 4), "cores": cores, "seeds": [1, 2, 3],
            "graph_recall_cpu_exact": [round(x, 5) for x in cpu_rec],
            "graph_recall_cpu_reference_sweep": [round(x, 5) for x in sweep_rec],
            "graph_recall_gpu": [round(x, 5) for x in gpu_rec],
            "max_abs_gap_gpu_vs_exact_pt": round(100 * max(abs(a_ - b_) for a_, b_ in zip(gpu_rec, cpu_rec)), 3),
            "gpu_equals_oracle_slot_mode_bit_exact": bool(np.array_equal(so, g1_first[:, :w1["K"]]))}
        # the 10k configuration's search: k = 10, L = 100, random entries
        b1q = getattr(datagen, w1["gen"])(**w1["gen_kw"])[1]
        gq1, _, _ = ctx.exact_knn(b1, b1q, 10)
        ix1 = Index(ctx, b1)
        ix1.build_nndescent(*prm, seed=1, dtype="f32")
        g1b, _, d1b = ix1.get_graph()
        s_ids, s_dist, _, _, s_st = ix1.search(b1q, 100, 10, seed=7)
        ix1.close()
        c0 = time.perf_counter()
        r_ids, r_dist, _, _ = oracle.search_greedy(g1b, d1b, b1, b1q, 100, 10, seed=7, nthreads=cores)
        c1_cpu_s = time.perf_counter() - c0
        out["cpu_baseline"]["c1_search"] = {
            "workload": "10k x 128, 1,000 queries, k=10, L=100, random entries, on the GPU's C1 graph (seed 1)",
            "gpu_qps": round(len(b1q) / (s_st["kernel_ms"] / 1e3)), "cpu_qps": round(len(b1q) / c1_cpu_s),
            "cores": cores, "recall_at_10": round(recall_at_k(s_ids, gq1, 10), 4),
            "ids_bit_exact_vs_oracle": bool(np.array_equal(s_ids, r_ids) and np.array_equal(s_dist, r_dist))}
        log(f"cpu baseline single thread {out['cpu_baseline']['single_thread_qps']} QPS; nndescent 10k "
            f"{out['cpu_baseline']['nndescent_10k']}")
    if w["learned"]:  # the 2M x 96 graph also serves the batching configuration
        out["batched_c4"] = batched_leg(ctx, ix, base, queries, dq, datagen, args)
    # roofline of the served method's launch: algorithmic bytes = rows
    # gathered for the reference's evaluation count (the kernel's count less
    # the re-evaluations of its forgetful visited table, measured against the
    # oracle's exact count on the CPU sample for plain search), adjacency
    # lists read, query and results
    ratio = out.get("cpu_baseline", {}).get("gpu_evals_over_reference_evals") if method.startswith("plain") else None
    ref_evals = evals / ratio if ratio else evals
    deg = w["K"]
    alg_bytes = 4 * d * ref_evals + 4 * deg * hops + Q * (4 * d + 8 * k)
    achieved = alg_bytes / (kern_ms / 1e3) / 1e9
    traffic = None
    tp = os.path.join(ROOT, "profiles", f"greedy_search_{args.workload}_traffic.json")
    if os.path.exists(tp):
        try:
            tj = json.load(open(tp))
            if tj.get("method") == method and tj.get("L") == L_plain:  # measured on this launch's shape
                traffic = tj.get("hbm_bytes_per_launch")
        except (ValueError, OSError):
            traffic = None
    out["roofline"] = {"bound": "hbm", "kernel": "greedy_search_kernel", "achieved": round(achieved, 1),
                       "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                       "traffic": traffic, "kernel_ms": round(kern_ms, 4), "alg_bytes_per_launch": int(alg_bytes),
                       "dist_evals_per_query": round(evals / Q, 1),
                       "reference_evals_per_query": round(ref_evals / Q, 1),
                       "hops_per_query": round(hops / Q, 1)}
    # An achievable ceiling for the access pattern: the search's distance
    # phase alone (random 64-row gathers of this base's rows, scored with the
    # search's gather tile) at the residency the served launch's LDS allows,
    # and at others for comparison.
    res = lds_residency(L_plain, plain_tuning)
    ceil = {}
    for wpc in sorted({4, res, 12, 16}):
        g_bps, _ = ctx.gather_rate(N, d, wpc, batches=256)
        ceil[str(wpc)] = round(g_bps, 1)
    out["roofline"]["gather_ceiling"] = {
        "unit": "GB/s", "by_waves_per_cu": ceil, "search_waves_per_cu": res,
        "frac_of_ceiling": round(achieved / ceil[str(res)], 4),
        "note": "random rows of the base gathered and scored 64 at a time with the search's tile, "
                "waves_per_cu wavefronts resident a CU (wv_gather_rate)"}
    log(f"gather ceiling {ceil} GB/s; search at {res} wavefronts/CU: {achieved:.0f} GB/s")
    print(json.dumps(out), flush=True)
    for a in (dq, d_ids, d_dist, d_ev, d_hp):
        a.free()
    ix.close()
    ctx.close()


if __name__ == "__main__":
    main()
