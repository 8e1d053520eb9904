"""wv_index_search_dev_noalloc against wv_index_search_dev, through the C ABI.

The no-allocation entry point keeps its per-query counts and statistics
totals in a fixed device array of the library.  Its results must equal the
allocating entry point's bit for bit (ids, distances, evaluation and hop
counts, statistics totals) on two calls in a row, the second with fewer
queries, so that nothing carries over from the first call's contents.
"""
import numpy as np
import pytest

import oracle as O
from weavess_amd import Index, WvError, datagen
from weavess_amd._capi import WV_EINVAL, WV_NOALLOC_MAX_Q
from weavess_amd.devmem import DeviceArray, pad_rows

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def index(ctx):
    base, q = datagen.small(3000, 2500, 128, clusters=12, seed=21, integer=False)
    ids, _, deg = O.nndescent(base, 32, 48, 10, 48, 8, seed=4)
    ix = Index(ctx, base)
    ix.set_graph(ids, deg)
    yield ix, q
    ix.close()


def _run(ix, q, L, k, noalloc, pass_counts):
    Q = len(q)
    dq = DeviceArray.from_host(pad_rows(q))
    di = DeviceArray((Q, k), np.uint32)
    dd = DeviceArray((Q, k), np.float32)
    de = DeviceArray((Q,), np.uint32) if pass_counts else None
    dh = DeviceArray((Q,), np.uint32) if pass_counts else None
    kw = dict(d_out_dist=dd.ptr, seed=11, d_evals=de.ptr if de else 0, d_hops=dh.ptr if dh else 0)
    if noalloc:
        st = ix.search_dev_noalloc(dq.ptr, Q, L, k, di.ptr, **kw)
    else:
        st = ix.search_dev(dq.ptr, Q, L, k, di.ptr, **kw)
    out = (di.to_host(), dd.to_host(), de.to_host() if de else None, dh.to_host() if dh else None, st)
    for a in (dq, di, dd, de, dh):
        if a is not None:
            a.free()
    return out


@pytest.mark.parametrize("pass_counts", [False, True])
def test_noalloc_equals_search_dev_two_calls(index, pass_counts):
    ix, q = index
    # first call 2,500 queries at L=64, then 700 queries at L=40
    for Q, L, k in ((2500, 64, 10), (700, 40, 5)):
        print(f"noalloc case Q={Q} L={L} k={k} pass_counts={pass_counts}", flush=True)
        a = _run(ix, q[:Q], L, k, False, pass_counts)
        b = _run(ix, q[:Q], L, k, True, pass_counts)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        if pass_counts:
            assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])
        assert a[4]["dist_evals"] == b[4]["dist_evals"] and a[4]["hops"] == b[4]["hops"]
        assert b[4]["hops"] >= Q and b[4]["kernel_ms"] > 0


def test_noalloc_rejects_too_many_queries(index):
    ix, _ = index
    with pytest.raises(WvError) as e:
        ix.search_dev_noalloc(1, WV_NOALLOC_MAX_Q + 1, 32, 10, 1)
    assert e.value.code == WV_EINVAL and "maximum" in str(e.value)
