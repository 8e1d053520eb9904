"""The C-ABI library (CPU only: load and symbol checks, no compute calls)."""
import ctypes
import os
import re

import pytest

from weavess_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "weavess_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wv_[a-z_0-9]+)\s*\(", text)))


def test_library_is_built_and_loads():
    assert os.path.exists(_capi.LIB_PATH), "run python -m weavess_amd.build"
    lib = _capi.load()
    assert lib.wv_version() == 1


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(_capi.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(_capi.EXPORTS) == syms


def test_struct_layouts_match_header():
    # sizes of the POD structs as laid out by the C compiler for the header
    assert ctypes.sizeof(_capi.NndParams) == 48
    assert ctypes.sizeof(_capi.BuildStats) == 80
    assert ctypes.sizeof(_capi.SearchStats) == 24
    assert ctypes.sizeof(_capi.Graph) == 24
    assert ctypes.sizeof(_capi.SearchConfig) == 24
    assert ctypes.sizeof(_capi.NsgParams) == 24
    assert ctypes.sizeof(_capi.NsgStats) == 56
    assert ctypes.sizeof(_capi.VamanaParams) == 32
    assert ctypes.sizeof(_capi.SsgParams) == 24


def test_null_context_is_rejected():
    lib = _capi.load()
    assert lib.wv_ctx_create(0, None) == _capi.WV_EINVAL
    assert lib.wv_ctx_destroy(None) == _capi.WV_EINVAL
    assert lib.wv_index_destroy(None) == _capi.WV_EINVAL
    assert lib.wv_last_error(None) == b"null context"
