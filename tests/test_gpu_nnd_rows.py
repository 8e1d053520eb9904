"""NN-Descent join at row lengths other than 128: a partial last MFMA step,
rows of several register pieces, and rows streamed through registers piece
by piece (longer than 512 bytes), against the CPU oracle."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def int_rows(n, d, hi, seed):
    """Clustered small integers in [0, hi]: exact in FP32 and BF16, and every
    norm, dot product and distance sum stays below 2^24."""
    rng = np.random.default_rng(seed)
    centres = rng.integers(0, hi + 1, size=(8, d))
    lab = rng.integers(0, 8, size=n)
    x = centres[lab] + rng.integers(-2, 3, size=(n, d))
    return np.clip(x, 0, hi).astype(np.float32)


@pytest.mark.parametrize("d, dtype", [
    (100, "f32"),   # 7 steps of 16, the last one partial (elements 96..99)
    (200, "f32"),   # 13 steps: longer than 512 bytes, streamed in k pieces of 2
    (136, "bf16"),  # 5 steps of 32 in the 8-step register tile
    (300, "bf16"),  # 10 steps: streamed
])
def test_nndescent_bit_exact_row_lengths(ctx, d, dtype):
    base = int_rows(600, d, 15, seed=d)
    K, L, S, R, it = 16, 24, 8, 24, 5
    gi, gd, gdeg, st = ctx.nndescent_build(base, K, L, S, R, it, seed=4, dtype=dtype, cand_slots=64)
    ri, rd, rdeg, rst = O.nndescent(base, K, L, S, R, it, seed=4, cand_slots=64, with_stats=True)
    assert st["pairs"] == sum(p for p, _ in rst)
    assert st["proposals"] == sum(a for _, a in rst)
    assert np.array_equal(gdeg, rdeg)
    assert np.array_equal(gi, ri)
    assert np.array_equal(gd, rd)


@pytest.mark.parametrize("d, dtype", [(200, "f32"), (300, "bf16"), (960, "f32"), (960, "bf16")])
def test_nndescent_bit_exact_long_lists(ctx, d, dtype):
    """Streamed rows with new lists of up to 56 (four row tiles: two blocks
    of row tiles) and up to 136 listed rows (nine column tiles: column
    blocks after the first), including d = 960 as in the precision
    configuration."""
    base = int_rows(900, d, 15, seed=d + 1)
    K, L, S, R, it = 24, 40, 16, 40, 5
    gi, gd, gdeg, st = ctx.nndescent_build(base, K, L, S, R, it, seed=6, dtype=dtype, cand_slots=64)
    ri, rd, rdeg, rst = O.nndescent(base, K, L, S, R, it, seed=6, cand_slots=64, with_stats=True)
    assert st["pairs"] == sum(p for p, _ in rst)
    assert st["proposals"] == sum(a for _, a in rst)
    assert np.array_equal(gdeg, rdeg)
    assert np.array_equal(gi, ri)
    assert np.array_equal(gd, rd)


@pytest.mark.parametrize("d, dtype", [(300, "f32"), (300, "bf16"), (600, "fp8")])
def test_nndescent_streamed_rows_quality(ctx, d, dtype):
    """Continuous rows longer than one register tile in each precision: the
    graph recall against exact kNN lands within the precision's tolerance of
    the oracle's FP32 build.  Five rounds over a noisy mixture stop short of
    convergence (oracle recall about 0.87), so a join that loses part of a
    row would show."""
    rng = np.random.default_rng(d)
    centres = rng.normal(size=(30, d)).astype(np.float32)
    base = (centres[rng.integers(0, 30, size=1200)] + 4.0 * rng.normal(size=(1200, d))).astype(np.float32)
    K, L, S, R, it = 16, 32, 10, 32, 5
    ids, dist, deg, st = ctx.nndescent_build(base, K, L, S, R, it, seed=5, dtype=dtype)
    gt, _ = O.exact_knn(base, base, 10, self_offset=0)
    ri, rd, rdeg = O.nndescent(base, K, L, S, R, it, seed=5)
    q_gpu = O.graph_quality(ids, deg, gt)
    q_ref = O.graph_quality(ri, rdeg, gt)
    print(d, dtype, "graph recall gpu", q_gpu, "oracle", q_ref)
    assert q_gpu >= q_ref - (0.005 if dtype == "f32" else 0.02)
    for i in range(0, len(base), 97):  # lists: distinct, no self, ascending FP32 distances
        row = ids[i, :deg[i]].tolist()
        assert i not in row and len(set(row)) == len(row)
        assert list(dist[i, :deg[i]]) == sorted(dist[i, :deg[i]])
