"""weavess_amd -- MI355X-native kNN-graph hot path of WEAVESS.

NN-Descent graph construction and greedy best-first search as hand-written
HIP kernels for gfx950 behind a C ABI (include/weavess_hip.h,
libweavess_hip.so).  This package holds the kernels (csrc/), the in-tree
build (build.py), the ctypes binding (_capi.py), seeded synthetic workloads
(datagen.py), graph/vector file formats (io.py) and the C++ pipeline layer
that mirrors WEAVESS's Component/IndexBuilder interface (host/).
"""
from ._capi import Context, Index, WvError, load  # noqa: F401

__all__ = ["Context", "Index", "WvError", "load"]
