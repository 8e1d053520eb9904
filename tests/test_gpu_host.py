"""The C++ pipeline (IndexBuilder + HIP stages) end to end on the GPU."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from weavess_amd import io

pytestmark = pytest.mark.gpu
HOST = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "weavess_amd", "host")


def _run(args, env):
    r = subprocess.run([os.path.join(HOST, "weavess_main"), *args], capture_output=True, text=True,
                       env={**os.environ, **env}, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_pipeline_build_search_matches_c_abi(ctx, tmp_path, small_int):
    base, q = small_int
    pre = str(tmp_path / "c1")
    io.write_fvecs(pre + "_base.fvecs", base)
    io.write_fvecs(pre + "_query.fvecs", q)
    gt, _ = O.exact_knn(base, q, 10)
    io.write_ivecs(pre + "_groundtruth.ivecs", gt)
    env = dict(WEAVESS_K="24", WEAVESS_L="32", WEAVESS_ITER="6", WEAVESS_S="10", WEAVESS_R="32",
               WEAVESS_SEED="3")
    out = _run(["kgraph", pre, "build"], env)
    assert "Build cost" in out
    gids, gdeg = io.read_graph(pre + ".kgraph.graph")
    # init (INIT_RANDOM on the GPU) + refine as two stages == one-call build with the same seed
    ids, dist, deg, _ = ctx.nndescent_build(base, 24, 32, 10, 32, 6, seed=3)
    assert np.array_equal(gdeg, deg) and np.array_equal(gids[:, :24], ids)
    out = _run(["kgraph", pre, "search", "40"], env)
    rec = float(re.search(r"10 NN accuracy: ([0-9.]+)", out).group(1))
    res, *_ = O.search_greedy(ids, deg, base, q, 40, 10, seed=3)
    assert rec == pytest.approx(O.recall(res, gt, 10), abs=1e-6)
    out = _run(["kgraph", pre, "search"], {**env, "WEAVESS_ACC_SET": "0.9"})
    m = re.search(r"QPS: ([0-9.e+]+) recall: ([0-9.]+) L: (\d+)", out)
    assert float(m.group(2)) >= 0.9


def test_pipeline_debug_monitor(ctx, tmp_path, small_int):
    """WEAVESS_DEBUG=1 passes refine's debug flag: the build prints the graph
    recall of 100 control nodes after every round (the reference's
    "NN-Descent iter" / "Graph Quality" lines) and builds the same graph."""
    base, q = small_int
    pre = str(tmp_path / "dbg")
    io.write_fvecs(pre + "_base.fvecs", base)
    io.write_fvecs(pre + "_query.fvecs", q)
    gt, _ = O.exact_knn(base, q, 10)
    io.write_ivecs(pre + "_groundtruth.ivecs", gt)
    env = dict(WEAVESS_K="24", WEAVESS_L="32", WEAVESS_ITER="6", WEAVESS_S="10", WEAVESS_R="32",
               WEAVESS_SEED="3", WEAVESS_DEBUG="1")
    out = _run(["kgraph", pre, "build"], env)
    q_lines = [float(v) for v in re.findall(r"Graph Quality : ([0-9.e+-]+)", out)]
    assert len(q_lines) == 6 and re.findall(r"NN-Descent iter: (\d+)", out) == [str(i) for i in range(6)]
    assert 0.0 < q_lines[0] < q_lines[-1] <= 0.32  # pools hold at most L = 32 of the top 100
    gids, gdeg = io.read_graph(pre + ".kgraph.graph")
    ids, _, deg, _ = ctx.nndescent_build(base, 24, 32, 10, 32, 6, seed=3)
    assert np.array_equal(gdeg, deg) and np.array_equal(gids[:, :24], ids)


def test_pipeline_centroid_entry(ctx, tmp_path, small_int):
    """WEAVESS_ENTRY=centroid: ep is the exact nearest base point to the
    centroid, and search from ep's neighbours plus the counter-stream fill
    gives the oracle's recall with the same explicit entry sets."""
    from weavess_amd.entries import centroid, centroid_entries, centroid_ep
    base, q = small_int
    pre = str(tmp_path / "c1")
    io.write_fvecs(pre + "_base.fvecs", base)
    io.write_fvecs(pre + "_query.fvecs", q)
    gt, _ = O.exact_knn(base, q, 10)
    io.write_ivecs(pre + "_groundtruth.ivecs", gt)
    env = dict(WEAVESS_K="24", WEAVESS_L="32", WEAVESS_ITER="6", WEAVESS_S="10", WEAVESS_R="32",
               WEAVESS_SEED="5", WEAVESS_ENTRY="centroid")
    _run(["kgraph", pre, "build"], env)
    gids, gdeg = io.read_graph(pre + ".kgraph.graph")
    out = _run(["kgraph", pre, "search", "40"], env)
    ep = int(re.search(r"ep_ (\d+)", out).group(1))
    d2 = ((base.astype(np.float64) - centroid(base).astype(np.float64)) ** 2).sum(axis=1)
    assert d2[ep] == d2.min()
    assert ep == centroid_ep(ctx, base)
    ent = centroid_entries(ep, gids, gdeg, len(q), 40, len(base), seed=5)
    assert np.array_equal(ent[:, :min(40, int(gdeg[ep]))],
                          np.broadcast_to(gids[ep, :min(40, int(gdeg[ep]))], (len(q), min(40, int(gdeg[ep])))))
    res, *_ = O.search_greedy(gids, gdeg, base, q, 40, 10, entries=ent)
    rec = float(re.search(r"10 NN accuracy: ([0-9.]+)", out).group(1))
    assert rec == pytest.approx(O.recall(res, gt, 10), abs=1e-6)
    # and the C ABI with the same entries returns the oracle's ids
    from weavess_amd import Index
    ix = Index(ctx, base)
    ix.set_graph(gids, gdeg)
    gres, *_ = ix.search(q, 40, 10, entries=ent)
    ix.close()
    assert np.array_equal(gres, res)


def test_pipeline_nsg_and_ssg_entries(ctx, tmp_path, small_int):
    """A build written as an NSG file carries the centroid entry point and a
    search of it reuses that ep; an SSG file's entry set eps leads every
    query's entries (SubCentroid).  Both give the oracle's recall with the
    same explicit entry sets."""
    from weavess_amd.entries import centroid_entries, centroid_ep, subcentroid_entries
    base, q = small_int
    pre = str(tmp_path / "c1")
    io.write_fvecs(pre + "_base.fvecs", base)
    io.write_fvecs(pre + "_query.fvecs", q)
    gt, _ = O.exact_knn(base, q, 10)
    io.write_ivecs(pre + "_groundtruth.ivecs", gt)
    gn = pre + ".nsg.graph"
    env = dict(WEAVESS_K="24", WEAVESS_L="32", WEAVESS_ITER="6", WEAVESS_S="10", WEAVESS_R="32",
               WEAVESS_SEED="9", WEAVESS_GRAPH=gn, WEAVESS_FORMAT="nsg")
    _run(["kgraph", pre, "build"], env)
    gids, gdeg, ep = io.read_graph(gn, fmt="nsg")
    assert ep == centroid_ep(ctx, base)
    out = _run(["kgraph", pre, "search", "48"], {**env, "WEAVESS_ENTRY": "centroid"})
    assert "__ENTRY : CENTROID__" not in out  # ep came from the file
    ent = centroid_entries(ep, gids, gdeg, len(q), 48, len(base), seed=9)
    res, *_ = O.search_greedy(gids, gdeg, base, q, 48, 10, entries=ent)
    assert float(re.search(r"10 NN accuracy: ([0-9.]+)", out).group(1)) == pytest.approx(
        O.recall(res, gt, 10), abs=1e-6)

    gs = pre + ".ssg.graph"
    eps = np.random.default_rng(4).choice(len(base), 8, replace=False).astype(np.uint32)
    io.write_graph(gs, gids, gdeg, fmt="ssg", ep=eps)
    senv = {**env, "WEAVESS_GRAPH": gs, "WEAVESS_FORMAT": "ssg", "WEAVESS_ENTRY": "subcentroid"}
    out = _run(["kgraph", pre, "search", "32"], senv)
    ent = subcentroid_entries(eps, len(q), 32, len(base), seed=9)
    res, *_ = O.search_greedy(gids, gdeg, base, q, 32, 10, entries=ent)
    assert float(re.search(r"10 NN accuracy: ([0-9.]+)", out).group(1)) == pytest.approx(
        O.recall(res, gt, 10), abs=1e-6)
    # SET_RECALL's L sweep runs with the same entry component
    out = _run(["kgraph", pre, "search"], {**senv, "WEAVESS_ACC_SET": "0.9"})
    assert float(re.search(r"recall: ([0-9.]+)", out).group(1)) >= 0.9
    # more entry points than L_search is an error, as the reference asserts
    io.write_graph(gs, gids, gdeg, fmt="ssg", ep=np.arange(12, dtype=np.uint32))
    r = subprocess.run([os.path.join(HOST, "weavess_main"), "kgraph", pre, "search", "10"], capture_output=True,
                       text=True, env={**os.environ, **senv}, timeout=300)
    assert r.returncode == 2 and "more entry points" in r.stderr


def test_pipeline_efanna_build_matches_oracle(ctx, tmp_path, small_int):
    """weavess_main efanna: INIT_RANDOM on the GPU, then REFINE_EFANNA from
    those rows, equals the oracle's EFANNA build from InitRandom bit for bit
    (integer data), and its search gives the oracle's recall."""
    base, q = small_int
    pre = str(tmp_path / "c1")
    io.write_fvecs(pre + "_base.fvecs", base)
    io.write_fvecs(pre + "_query.fvecs", q)
    gt, _ = O.exact_knn(base, q, 10)
    io.write_ivecs(pre + "_groundtruth.ivecs", gt)
    env = dict(WEAVESS_K="24", WEAVESS_L="32", WEAVESS_ITER="4", WEAVESS_S="10", WEAVESS_R="32",
               WEAVESS_SEED="6")
    out = _run(["efanna", pre, "build"], env)
    assert "__REFINE : EFANNA (HIP)__" in out
    gids, gdeg = io.read_graph(pre + ".efanna.graph")
    ri, _, rdeg = O.nndescent_efanna(base, 24, 32, 10, 32, 4, seed=6, cand_slots=128)
    assert np.array_equal(gdeg, rdeg) and np.array_equal(gids[:, :24], ri)
    out = _run(["efanna", pre, "search", "40"], env)
    res, *_ = O.search_greedy(ri, rdeg, base, q, 40, 10, seed=6)
    assert float(re.search(r"10 NN accuracy: ([0-9.]+)", out).group(1)) == pytest.approx(
        O.recall(res, gt, 10), abs=1e-6)


def test_pipeline_nsg_build_and_search(ctx, tmp_path, small_int):
    """weavess_main nsg: NN-Descent, the NSG refine from the centroid entry
    point, the NSG file layout; equal to the oracle's NSG of the oracle's
    kNN graph, and its centroid-entry search gives the oracle's recall."""
    from weavess_amd.entries import centroid_entries, centroid_ep
    base, q = small_int
    pre = str(tmp_path / "c1")
    io.write_fvecs(pre + "_base.fvecs", base)
    io.write_fvecs(pre + "_query.fvecs", q)
    gt, _ = O.exact_knn(base, q, 10)
    io.write_ivecs(pre + "_groundtruth.ivecs", gt)
    env = dict(WEAVESS_K="24", WEAVESS_L="32", WEAVESS_ITER="6", WEAVESS_S="10", WEAVESS_R="32",
               WEAVESS_SEED="4", WEAVESS_L_REFINE="40", WEAVESS_R_REFINE="16", WEAVESS_C_REFINE="200")
    out = _run(["nsg", pre, "build"], env)
    assert "__REFINE : NSG (HIP)__" in out
    gids, gdeg, ep = io.read_graph(pre + ".nsg.graph", fmt="nsg")
    assert ep == centroid_ep(ctx, base)
    kid, _, kdeg = O.nndescent(base, 24, 32, 10, 32, 6, seed=4, cand_slots=128)
    oi, _, odeg, _ = O.nsg(base, kid, kdeg, 40, 16, 200, ep, seed=4)
    assert np.array_equal(gdeg, odeg)
    for i in range(len(base)):
        assert gids[i, :gdeg[i]].tolist() == oi[i, :odeg[i]].tolist(), i
    out = _run(["nsg", pre, "search", "20"], env)
    ent = centroid_entries(ep, oi, odeg, len(q), 20, len(base), seed=4)
    res, *_ = O.search_greedy(oi, odeg, base, q, 20, 10, entries=ent)
    assert float(re.search(r"10 NN accuracy: ([0-9.]+)", out).group(1)) == pytest.approx(
        O.recall(res, gt, 10), abs=1e-6)


def test_pipeline_vamana_build_and_search(ctx, tmp_path, small_int):
    """weavess_main vamana: INIT_RAND, two Vamana passes (alpha 1, 2), the
    Vamana layout with ep; equal to two oracle passes from the oracle's
    random graph, and its centroid search gives the oracle's recall."""
    from weavess_amd.entries import centroid_entries, centroid_ep
    base, q = small_int
    pre = str(tmp_path / "c1")
    io.write_fvecs(pre + "_base.fvecs", base)
    io.write_fvecs(pre + "_query.fvecs", q)
    gt, _ = O.exact_knn(base, q, 10)
    io.write_ivecs(pre + "_groundtruth.ivecs", gt)
    env = dict(WEAVESS_SEED="5", WEAVESS_L_REFINE="40", WEAVESS_R_REFINE="24", WEAVESS_L="24",
               WEAVESS_C_REFINE="300")
    out = _run(["vamana", pre, "build"], env)
    assert out.count("__REFINE : VAMANA (HIP)__") == 2
    gids, gdeg, ep = io.read_graph(pre + ".vamana.graph", fmt="nsg")
    assert ep == centroid_ep(ctx, base)
    rid, _, rcnt = O.init_random(base, 24, seed=5)
    g1, _, d1 = O.vamana(base, rid, rcnt, 40, 24, 300, 1.0, ep, seed=5)
    g2, _, d2 = O.vamana(base, g1, d1, 40, 24, 300, 2.0, ep, seed=5)
    assert np.array_equal(gdeg, d2)
    for i in range(len(base)):
        assert gids[i, :gdeg[i]].tolist() == g2[i, :d2[i]].tolist(), i
    out = _run(["vamana", pre, "search", "60"], env)
    ent = centroid_entries(ep, g2, d2, len(q), 60, len(base), seed=5)
    res, *_ = O.search_greedy(g2, d2, base, q, 60, 10, entries=ent)
    assert float(re.search(r"10 NN accuracy: ([0-9.]+)", out).group(1)) == pytest.approx(
        O.recall(res, gt, 10), abs=1e-6)


def test_pipeline_dpg_build_and_search(ctx, tmp_path, small_int):
    """weavess_main dpg: NN-Descent, then DPG (K / 2 a node, mirrored);
    equal to the oracle's DPG of the oracle's kNN graph, and the search
    from random entries gives the oracle's recall."""
    base, q = small_int
    pre = str(tmp_path / "c1")
    io.write_fvecs(pre + "_base.fvecs", base)
    io.write_fvecs(pre + "_query.fvecs", q)
    gt, _ = O.exact_knn(base, q, 10)
    io.write_ivecs(pre + "_groundtruth.ivecs", gt)
    env = dict(WEAVESS_K="24", WEAVESS_L="32", WEAVESS_ITER="6", WEAVESS_S="10", WEAVESS_R="32",
               WEAVESS_SEED="7")
    out = _run(["dpg", pre, "build"], env)
    assert "__REFINE : DPG (HIP)__" in out
    gids, gdeg = io.read_graph(pre + ".dpg.graph")
    kid, _, kdeg = O.nndescent(base, 24, 32, 10, 32, 6, seed=7, cand_slots=128)
    oi, _, odeg = O.dpg(base, kid, kdeg, 12)
    assert np.array_equal(gdeg, odeg)
    for i in range(len(base)):
        assert gids[i, :gdeg[i]].tolist() == oi[i, :odeg[i]].tolist(), i
    out = _run(["dpg", pre, "search", "30"], env)
    res, *_ = O.search_greedy(oi, odeg, base, q, 30, 10, seed=7)
    assert float(re.search(r"10 NN accuracy: ([0-9.]+)", out).group(1)) == pytest.approx(
        O.recall(res, gt, 10), abs=1e-6)


def test_pipeline_ssg_build_and_search(ctx, tmp_path, small_int):
    """weavess_main ssg: NN-Descent, then SSG (roots eps_ in the SSG
    layout); equal to the oracle's SSG of the oracle's kNN graph, and the
    sub-centroid search from eps_ gives the oracle's recall."""
    from weavess_amd.entries import subcentroid_entries
    base, q = small_int
    pre = str(tmp_path / "c1")
    io.write_fvecs(pre + "_base.fvecs", base)
    io.write_fvecs(pre + "_query.fvecs", q)
    gt, _ = O.exact_knn(base, q, 10)
    io.write_ivecs(pre + "_groundtruth.ivecs", gt)
    env = dict(WEAVESS_K="24", WEAVESS_L="32", WEAVESS_ITER="6", WEAVESS_S="10", WEAVESS_R="32",
               WEAVESS_SEED="8", WEAVESS_L_REFINE="50", WEAVESS_R_REFINE="20")
    out = _run(["ssg", pre, "build"], env)
    assert "__REFINE : SSG (HIP)__" in out
    gids, gdeg, eps = io.read_graph(pre + ".ssg.graph", fmt="ssg")
    kid, _, kdeg = O.nndescent(base, 24, 32, 10, 32, 6, seed=8, cand_slots=128)
    oi, _, odeg, oeps, _ = O.ssg(base, kid, kdeg, 50, 20, 60.0, 10, seed=8)
    assert eps.tolist() == oeps.tolist() and np.array_equal(gdeg, odeg)
    for i in range(len(base)):
        assert gids[i, :gdeg[i]].tolist() == oi[i, :odeg[i]].tolist(), i
    out = _run(["ssg", pre, "search", "30"], env)
    ent = subcentroid_entries(oeps, len(q), 30, len(base), seed=8)
    res, *_ = O.search_greedy(oi, odeg, base, q, 30, 10, entries=ent)
    assert float(re.search(r"10 NN accuracy: ([0-9.]+)", out).group(1)) == pytest.approx(
        O.recall(res, gt, 10), abs=1e-6)
