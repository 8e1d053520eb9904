"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def ctx():
    """One device context for the GPU tests (fails loudly without the library)."""
    from weavess_amd import Context
    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def small_int():
    """Integer-valued 2,000 x 128 base + 200 queries (C1-like mixture)."""
    from weavess_amd import datagen
    return datagen.small(2000, 200, 128, clusters=10, seed=7, integer=True)


@pytest.fixture(scope="session")
def small_float():
    """Continuous 2,000 x 128 base + 200 queries."""
    from weavess_amd import datagen
    return datagen.small(2000, 200, 128, clusters=10, seed=8, integer=False)


@pytest.fixture(scope="session")
def small_int_graph(small_int):
    """Oracle NN-Descent graph (K=32) of small_int."""
    import oracle as O
    base, _ = small_int
    ids, dist, deg = O.nndescent(base, 32, 48, 10, 48, 8, seed=11)
    return ids, dist, deg


@pytest.fixture(scope="session")
def small_float_graph(small_float):
    import oracle as O
    base, _ = small_float
    ids, dist, deg = O.nndescent(base, 32, 48, 10, 48, 8, seed=12)
    return ids, dist, deg


def chain_points(n=8, d=4):
    """Points (i, 0, 0, 0), i = 0..n-1."""
    p = np.zeros((n, d), np.float32)
    p[:, 0] = np.arange(n)
    return p
