"""EFANNA's NN-Descent on the GPU vs the CPU oracle."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _build(ctx, base, *args, **kw):
    from weavess_amd import Index
    ix = Index(ctx, base)
    try:
        st = ix.build_efanna(*args, **kw)
        ids, dist, deg = ix.get_graph()
    finally:
        ix.close()
    return ids, dist, deg, st


def _init_rows(N, stride, seed):
    """A ragged initial graph with self, repeated and out-of-range ids."""
    rng = np.random.default_rng(seed)
    rows = rng.integers(0, N, size=(N, stride)).astype(np.uint32)
    rows[::3, 0] = np.arange(0, N, 3)
    rows[::4, 1] = rows[::4, 2]
    rows[::5, 3] = 0xFFFFFFF0
    cnt = rng.integers(1, stride + 1, size=N).astype(np.uint32)
    cnt[:7] = stride
    return rows, cnt


@pytest.mark.parametrize("slots", [64, 256])
def test_efanna_bit_exact_random_init(ctx, small_int, slots):
    """Integer data: every FP32 distance is exact, so the GPU build equals
    the oracle's (same streams, same candidate slots) bit for bit."""
    base, _ = small_int
    K, L, S, R, it = 24, 32, 10, 32, 5
    print("efanna random init, slots", slots, flush=True)
    gi, gd, gdeg, st = _build(ctx, base, K, L, S, R, it, seed=3, cand_slots=slots)
    ri, rd, rdeg, rst = O.nndescent_efanna(base, K, L, S, R, it, seed=3, cand_slots=slots, with_stats=True)
    assert st["pairs"] == sum(p for p, _ in rst)
    assert st["proposals"] == sum(a for _, a in rst)
    assert np.array_equal(gdeg, rdeg) and np.array_equal(gi, ri) and np.array_equal(gd, rd)


@pytest.mark.parametrize("stride,L,S,it", [(96, 128, 12, 0), (96, 128, 12, 3), (40, 40, 12, 3), (60, 80, 35, 3)])
def test_efanna_bit_exact_initial_graph(ctx, small_int, stride, L, S, it):
    """Initial rows wider than a wavefront, ragged, with self / repeats /
    out-of-range ids; ITER=0 returns them sorted and cleaned; S=35 draws
    70 first-round samples (more than one wavefront)."""
    base, _ = small_int
    rows, cnt = _init_rows(len(base), stride, 9)
    K, R = 24, 32
    print("efanna init graph", stride, L, S, it, flush=True)
    gi, gd, gdeg, _ = _build(ctx, base, K, L, S, R, it, seed=4, cand_slots=256, init_ids=rows, init_cnt=cnt)
    ri, rd, rdeg = O.nndescent_efanna(base, K, L, S, R, it, seed=4, cand_slots=256, init=(rows, cnt))
    assert np.array_equal(gdeg, rdeg) and np.array_equal(gi, ri) and np.array_equal(gd, rd)


@pytest.mark.parametrize("d, dtype", [(200, "f32"), (300, "bf16")])
def test_efanna_bit_exact_streamed_rows(ctx, d, dtype):
    """Rows longer than 512 bytes (the join's streamed form) with S=35:
    the first round's new lists hold 70 samples and more (five row tiles).
    Small integers keep every FP32 and BF16 quantity exact."""
    rng = np.random.default_rng(d)
    centres = rng.integers(0, 16, size=(8, d))
    base = np.clip(centres[rng.integers(0, 8, size=700)] + rng.integers(-2, 3, size=(700, d)), 0, 15)
    base = base.astype(np.float32)
    K, L, S, R, it = 24, 64, 35, 32, 4
    print("efanna streamed rows", d, dtype, flush=True)
    gi, gd, gdeg, st = _build(ctx, base, K, L, S, R, it, seed=7, dtype=dtype, cand_slots=128)
    ri, rd, rdeg, rst = O.nndescent_efanna(base, K, L, S, R, it, seed=7, cand_slots=128, with_stats=True)
    assert st["pairs"] == sum(p for p, _ in rst)
    assert st["proposals"] == sum(a for _, a in rst)
    assert np.array_equal(gdeg, rdeg) and np.array_equal(gi, ri) and np.array_equal(gd, rd)


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
def test_efanna_quality(ctx, small_float, dtype):
    base, _ = small_float
    K, L, S, R, it = 20, 32, 10, 32, 6
    ids, dist, deg, st = _build(ctx, base, K, L, S, R, it, seed=5, dtype=dtype)
    gt, _ = O.exact_knn(base, base, 10, self_offset=0)
    q_gpu = O.graph_quality(ids, deg, gt)
    ri, _, rdeg = O.nndescent_efanna(base, K, L, S, R, it, seed=5)
    q_ref = O.graph_quality(ri, rdeg, gt)
    print(dtype, "graph recall gpu", q_gpu, "oracle", q_ref, st)
    assert q_gpu >= q_ref - (0.005 if dtype == "f32" else 0.02)
    for i in range(0, len(base), 97):
        row = ids[i, :deg[i]].tolist()
        assert i not in row and len(set(row)) == len(row)


def test_efanna_rejects_bad_params(ctx, small_float):
    from weavess_amd import WvError
    base, _ = small_float
    with pytest.raises(WvError):
        _build(ctx, base, 16, 80, 65, 16, 2)                      # S above 64
    with pytest.raises(WvError):
        _build(ctx, base[:20], 4, 16, 10, 16, 2)                  # 2S >= N
    with pytest.raises(WvError):
        _build(ctx, base, 16, 32, 8, 16, 2, init_ids=np.zeros((len(base), 33), np.uint32))  # stride > L
