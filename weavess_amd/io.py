"""File formats of the reference pipeline.

fvecs / ivecs (src/component_load.cpp:9-28): per row an int32 dimension
followed by that many float32 / uint32 values.

Graph files (src/builder.cpp:1068-1076 writer, :1301-1308 reader): per node
a uint32 degree followed by that many uint32 neighbour ids; no distances.
KGraph files have no header.  NSG and Vamana files start with the entry
point ep (one uint32), SSG files with n_ep and then n_ep entry ids
(src/builder.cpp:920-925 writer, :1089-1095 reader).  ``read_graph``
returns the padded adjacency form the C ABI takes (ids N x deg_max, padded
with 0xFFFFFFFF, and deg[N]), plus the header for fmt "nsg" / "ssg".
"""
from __future__ import annotations

import numpy as np

PAD = 0xFFFFFFFF


def write_vecs(path: str, a: np.ndarray) -> None:
    a = np.ascontiguousarray(a)
    if a.dtype not in (np.float32, np.uint32, np.int32):
        raise TypeError("fvecs/ivecs hold 4-byte values")
    n, d = a.shape
    rows = np.empty((n, d + 1), np.uint32)
    rows[:, 0] = d
    rows[:, 1:] = a.view(np.uint32)
    rows.tofile(path)


def read_vecs(path: str, dtype=np.float32) -> np.ndarray:
    raw = np.fromfile(path, np.uint32)
    if raw.size == 0:
        return np.zeros((0, 0), dtype)
    d = int(raw[0])
    if d <= 0 or raw.size % (d + 1):
        raise ValueError(f"{path}: not a vecs file (dim {d}, {raw.size} words)")
    rows = raw.reshape(-1, d + 1)
    if np.any(rows[:, 0] != d):
        raise ValueError(f"{path}: rows of differing dimension")
    return np.ascontiguousarray(rows[:, 1:]).view(dtype)


def write_fvecs(path, a):
    write_vecs(path, np.asarray(a, np.float32))


def write_ivecs(path, a):
    write_vecs(path, np.asarray(a, np.uint32))


def read_fvecs(path):
    return read_vecs(path, np.float32)


def read_ivecs(path):
    return read_vecs(path, np.uint32)


_FORMATS = ("kgraph", "nsg", "ssg")


def write_graph(path: str, ids: np.ndarray, deg: np.ndarray, fmt: str = "kgraph", ep=None) -> None:
    """`ep`: the entry point (fmt "nsg") or the list of entry ids (fmt "ssg")."""
    if fmt not in _FORMATS:
        raise ValueError(f"unknown graph format {fmt!r}")
    if fmt != "kgraph" and ep is None:
        raise ValueError(f"a {fmt} graph file needs its entry point(s)")
    ids = np.asarray(ids, np.uint32)
    deg = np.asarray(deg, np.uint32)
    parts = []
    if fmt == "nsg":
        parts.append(np.array([ep], np.uint32))
    elif fmt == "ssg":
        eps = np.asarray(ep, np.uint32).ravel()
        parts += [np.array([eps.size], np.uint32), eps]
    for i in range(ids.shape[0]):
        parts.append(np.array([deg[i]], np.uint32))
        parts.append(ids[i, :deg[i]])
    (np.concatenate(parts) if parts else np.zeros(0, np.uint32)).tofile(path)


def read_graph(path: str, fmt: str = "kgraph"):
    if fmt not in _FORMATS:
        raise ValueError(f"unknown graph format {fmt!r}")
    raw = np.fromfile(path, np.uint32)
    degs, starts = [], []
    pos = 0
    head = None
    if fmt == "nsg":
        if raw.size < 1:
            raise ValueError(f"{path}: truncated graph header")
        head, pos = int(raw[0]), 1
    elif fmt == "ssg":
        if raw.size < 1 or raw.size < 1 + int(raw[0]):
            raise ValueError(f"{path}: truncated graph header")
        pos = 1 + int(raw[0])
        head = raw[1:pos].copy()
    while pos < raw.size:
        g = int(raw[pos])
        if pos + 1 + g > raw.size:
            raise ValueError(f"{path}: truncated graph file")
        degs.append(g)
        starts.append(pos + 1)
        pos += 1 + g
    n = len(degs)
    dmax = max(degs) if degs else 1
    ids = np.full((n, max(dmax, 1)), PAD, np.uint32)
    for i, (s, g) in enumerate(zip(starts, degs)):
        ids[i, :g] = raw[s:s + g]
    if fmt == "kgraph":
        return ids, np.asarray(degs, np.uint32)
    return ids, np.asarray(degs, np.uint32), head
