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
