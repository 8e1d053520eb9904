"""File formats and the C++ pipeline layer's CPU-side behaviour."""
import os
import struct
import subprocess

import numpy as np
import pytest

from weavess_amd import io

HOST = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "weavess_amd", "host")


def test_vecs_byte_layout(tmp_path):
    p = str(tmp_path / "a.fvecs")
    io.write_fvecs(p, np.array([[1.5, -2.0], [3.0, 4.0]], np.float32))
    raw = open(p, "rb").read()
    assert raw[:12] == struct.pack("<iff", 2, 1.5, -2.0) and len(raw) == 24
    assert np.array_equal(io.read_fvecs(p), np.array([[1.5, -2.0], [3.0, 4.0]], np.float32))


def test_ivecs_round_trip(tmp_path):
    p = str(tmp_path / "g.ivecs")
    a = np.arange(30, dtype=np.uint32).reshape(3, 10)
    io.write_ivecs(p, a)
    assert np.array_equal(io.read_ivecs(p), a)


def test_graph_byte_layout_and_round_trip(tmp_path):
    p = str(tmp_path / "x.graph")
    ids = np.array([[5, 6, 7], [1, 0, 0], [2, 3, 0]], np.uint32)
    deg = np.array([3, 1, 0], np.uint32)
    io.write_graph(p, ids, deg)
    assert open(p, "rb").read() == struct.pack("<IIIIIII", 3, 5, 6, 7, 1, 1, 0)
    rids, rdeg = io.read_graph(p)
    assert rdeg.tolist() == [3, 1, 0] and rids[0].tolist() == [5, 6, 7] and rids[1, 0] == 1


def test_graph_truncated_file_is_an_error(tmp_path):
    p = str(tmp_path / "t.graph")
    np.array([4, 1, 2], np.uint32).tofile(p)
    with pytest.raises(ValueError):
        io.read_graph(p)


def test_host_driver_built_and_prints_usage():
    exe = os.path.join(HOST, "weavess_main")
    assert os.path.exists(exe), "run __graft_entry__.build()"
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and "kgraph" in r.stdout


def test_host_driver_missing_file_is_an_error(tmp_path):
    exe = os.path.join(HOST, "weavess_main")
    r = subprocess.run([exe, "kgraph", str(tmp_path / "none"), "build"], capture_output=True, text=True)
    assert r.returncode == 2 and "open file error" in r.stderr


def test_centroid_entries_rules():
    """ep's neighbours first (duplicates and padding ids skipped), then
    distinct fill ids; deterministic per (seed, query)."""
    from weavess_amd.entries import centroid_entries
    N = 50
    graph = np.array([[3, 3, 7, 99, 1, 0, 0, 0]] * N, np.uint32)
    deg = np.full(N, 6, np.uint32)
    e = centroid_entries(5, graph, deg, 4, 10, N, seed=2)
    assert e.shape == (4, 10)
    assert np.all(e[:, :4] == [3, 7, 1, 0])  # 3 once, 99 >= N skipped, degree 6 caps the row
    for row in e:
        assert len(set(row.tolist())) == 10 and row.max() < N
    assert np.array_equal(e, centroid_entries(5, graph, deg, 4, 10, N, seed=2))
    assert not np.array_equal(e[:, 4:], centroid_entries(5, graph, deg, 4, 10, N, seed=3)[:, 4:])
    short = centroid_entries(5, graph, deg, 1, 2, N)  # L below the degree: only ep's first neighbours
    assert short[0].tolist() == [3, 7]


def test_graph_headers_byte_layout(tmp_path):
    """NSG/Vamana files lead with ep, SSG files with n_ep and the eps ids."""
    ids = np.array([[5, 6], [1, 0]], np.uint32)
    deg = np.array([2, 1], np.uint32)
    p = str(tmp_path / "n.graph")
    io.write_graph(p, ids, deg, fmt="nsg", ep=9)
    assert open(p, "rb").read() == struct.pack("<IIIIII", 9, 2, 5, 6, 1, 1)
    rids, rdeg, ep = io.read_graph(p, fmt="nsg")
    assert ep == 9 and rdeg.tolist() == [2, 1] and rids[0].tolist() == [5, 6]
    io.write_graph(p, ids, deg, fmt="ssg", ep=[4, 0, 3])
    assert open(p, "rb").read() == struct.pack("<IIIIIIIII", 3, 4, 0, 3, 2, 5, 6, 1, 1)
    rids, rdeg, eps = io.read_graph(p, fmt="ssg")
    assert eps.tolist() == [4, 0, 3] and rdeg.tolist() == [2, 1]
    with pytest.raises(ValueError):
        io.write_graph(p, ids, deg, fmt="nsg")
    np.array([5, 1], np.uint32).tofile(p)
    with pytest.raises(ValueError):
        io.read_graph(p, fmt="ssg")


def test_host_driver_reads_graph_headers(tmp_path):
    """weavess_main info with WEAVESS_FORMAT skips the NSG / SSG header; the
    plain reader rejects the same file (the header misreads as a row)."""
    rng = np.random.default_rng(0)
    pre = str(tmp_path / "p")
    b = rng.standard_normal((50, 8)).astype(np.float32)
    io.write_fvecs(pre + "_base.fvecs", b)
    io.write_fvecs(pre + "_query.fvecs", b[:5])
    io.write_ivecs(pre + "_groundtruth.ivecs", np.zeros((5, 10), np.uint32))
    ids = np.tile(np.arange(4, dtype=np.uint32), (50, 1))
    deg = np.full(50, 4, np.uint32)
    deg[3] = 2
    gn, gs = str(tmp_path / "g.nsg"), str(tmp_path / "g.ssg")
    io.write_graph(gn, ids, deg, fmt="nsg", ep=7)
    io.write_graph(gs, ids, deg, fmt="ssg", ep=[1, 2, 3])
    exe = os.path.join(HOST, "weavess_main")

    def info(graph, fmt):
        return subprocess.run([exe, "kgraph", pre, "info"], capture_output=True, text=True,
                              env={**os.environ, "WEAVESS_GRAPH": graph, "WEAVESS_FORMAT": fmt})
    r = info(gn, "nsg")
    assert r.returncode == 0 and "edges: 198 max degree: 4 min degree: 2" in r.stdout and "ep: 7" in r.stdout
    r = info(gs, "ssg")
    assert r.returncode == 0 and "edges: 198" in r.stdout and "eps: 3" in r.stdout
    r = info(gn, "kgraph")
    assert r.returncode == 2 and "truncated graph file" in r.stderr
    r = subprocess.run([exe, "kgraph", pre, "build"], capture_output=True, text=True,
                       env={**os.environ, "WEAVESS_FORMAT": "ssg"})
    assert r.returncode == 2 and "SSG files are written by ssg builds" in r.stderr


def test_subcentroid_entries_rules():
    """eps first (deduplicated, ids >= N skipped), then distinct fill ids
    from a stream of its own; more entry points than L is an error."""
    from weavess_amd.entries import centroid_entries, subcentroid_entries
    N = 40
    e = subcentroid_entries([9, 2, 9, 77, 5], 3, 8, N, seed=1)
    assert np.all(e[:, :3] == [9, 2, 5])
    for row in e:
        assert len(set(row.tolist())) == 8 and row.max() < N
    graph = np.array([[9, 2, 5, 0]] * N, np.uint32)
    c = centroid_entries(0, graph, np.full(N, 3, np.uint32), 3, 8, N, seed=1)
    assert np.array_equal(c[:, :3], e[:, :3]) and not np.array_equal(c[:, 3:], e[:, 3:])
    with pytest.raises(ValueError):
        subcentroid_entries(list(range(9)), 1, 8, N)
    assert subcentroid_entries([], 2, 4, N).shape == (2, 4)
