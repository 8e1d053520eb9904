"""Hand-worked known-answer case for NN-Descent's update() and join().

Six points on a line, x = 0, 1, 3, 7, 12, 20 (ids 0..5, d = 1), so every
squared distance is a distinct integer:

    d(0,1)=1   d(0,2)=9   d(0,3)=49  d(0,4)=144 d(0,5)=400
    d(1,2)=4   d(1,3)=36  d(1,4)=121 d(1,5)=361
    d(2,3)=16  d(2,4)=81  d(2,5)=289
    d(3,4)=25  d(3,5)=169
    d(4,5)=64

S = 2, L = K = 3, R = 8 (R >= N, so no reverse cap fires), exact insertion,
one thread.  Initial graph (S ids per node, all flagged new, M = S):

    P0 = (1,1) (9,2)      P1 = (1,0) (4,2)      P2 = (4,1) (16,3)
    P3 = (49,0) (169,5)   P4 = (64,5) (144,0)   P5 = (169,3) (361,1)

Entries are written (dist, id); N = new (flagged), o = old.

Round 0 (src/component_refine.cpp:121-124).  join() is empty: nn_new holds
2S zeros (include/weavess/index.h:98-103).  update() (:145-239):
  * M: maxl = min(M + S, |pool|) = 2; the scan stops after 2 flagged
    entries, so M = 2 everywhere (:166-174).
  * worst W = 9, 4, 16, 169, 144, 361 (pool.back()).
  * forward samples: every entry is flagged, so it goes to nn_new and its
    flag is cleared (:184-189).
  * reverse test `dist > W[nb]` (:185, :196):
      n=0: 1 > W1=4 no, 9 > W2=16 no;   n=1: 1 > 9 no, 4 > 16 no;
      n=2: 4 > W1=4 NO (equal), 16 > 169 no;
      n=3: 49 > W0=9 YES -> rnn_new[0] += 3;  169 > W5=361 no;
      n=4: 64 > 361 no;  144 > W0=9 YES -> rnn_new[0] += 4;
      n=5: 169 > W3=169 NO (equal);  361 > W1=4 YES -> rnn_new[1] += 5.
  * nn_new += rnn_new: 0:[1,2,3,4] 1:[0,2,5] 2:[1,3] 3:[0,5] 4:[5,0]
    5:[3,1]; every nn_old empty.

Round 1.  join() per node n, pairs (i, j) with i < j over nn_new (no
nn_old); each pair inserts both ways (include/weavess/index.h:120-149):
  n=0: (1,2)=4 dup both; (1,3)=36 -> P3 push (36,1); (1,4)=121 -> P4 push
       (121,1); (2,3)=16 -> P3 replaces (169,5); (2,4)=81 -> P4 replaces
       (144,0); (3,4)=25 -> P3 replaces (49,0), P4 replaces (121,1);
  n=1: (0,2)=9 -> P2 push (9,0); (0,5)=400 rejected (> 9, > 361);
       (2,5)=289 -> P5 push (289,2);
  n=2: (1,3) dup/rejected;  n=3: (0,5) rejected;  n=4: (0,5) rejected;
  n=5: (1,3) rejected.
  13 pairs, 8 accepted inserts.  Pools:
    P0 (1,1)o (9,2)o   P1 (1,0)o (4,2)o   P2 (4,1)o (9,0)N (16,3)o
    P3 (16,2)N (25,4)N (36,1)N   P4 (25,3)N (64,5)o (81,2)N
    P5 (169,3)o (289,2)N (361,1)o
  update(): maxl = min(2 + 2, |pool|):
    M = 2, 2, 3 (o N o: window ends at maxl), 2 (two flagged reached),
        3 (N o N), 3 (o N o).
    W = 9, 4, 16, 36, 81, 361.
    reverse: n=4: 81 > W2=16 -> rnn_new[2] += 4;
             n=5: 169 > W3=36 -> rnn_old[3] += 5; 289 > 16 -> rnn_new[2]
             += 5; 361 > W1=4 -> rnn_old[1] += 5; all others no.
    (36,1) of P3 lies outside its window and stays flagged.
  Lists: nn_new 0:[] 1:[] 2:[0,4,5] 3:[2,4] 4:[3,2] 5:[2]
         nn_old 0:[1,2] 1:[0,2,5] 2:[1,3] 3:[5] 4:[5] 5:[3,1]

Round 2.  join(): 17 pairs; the only accepted insert is (4,5)=64 into P5,
replacing (361,1).  update(): M = 2, 2, 3, 3, 3, 3;
W = 9, 4, 16, 36, 81, 289; reverse: n=3: 36 > W1=4 -> rnn_new[1] += 3;
n=4: 81 > 16 -> rnn_old[2] += 4; n=5: 169 > 36 -> rnn_old[3] += 5,
289 > 16 -> rnn_old[2] += 5.
  Lists: nn_new 1:[3] 3:[1] 5:[4], others empty;
         nn_old 0:[1,2] 1:[0,2] 2:[1,0,3,4,5] 3:[2,4,5] 4:[3,5,2] 5:[3,2]

Round 3.  join(): 7 pairs, nothing accepted.

Final graphs (sort, first K) after ITER = 2 and 3 rounds.  Nodes 0 and 1
keep two neighbours: include/weavess/index.h:122 rejects a candidate worse
than the pool's worst even when the pool has room.

Mistakes these values catch: an inverted or non-strict reverse test
(round 0 changes), sampling every flagged entry or a window without the
+S (rounds 1 and 2 change M and the lists), flags not cleared, reverse
lists not appended.
"""
import numpy as np
import pytest

import oracle as O

X = np.array([0, 1, 3, 7, 12, 20], np.float32).reshape(6, 1)
INIT_IDS = np.array([[1, 2], [0, 2], [1, 3], [0, 5], [5, 0], [3, 1]], np.uint32)
INIT_DIST = ((X[INIT_IDS, 0] - X[:, :1]) ** 2).astype(np.float32)
INIT = (INIT_IDS, INIT_DIST, np.full(6, 2, np.uint32))

# (dist, id, flag) per node after ITER rounds, then M, nn_new, nn_old, and
# (pairs, inserts) per round
EXPECTED = {
    1: dict(
        pools=[[(1, 1, 0), (9, 2, 0)], [(1, 0, 0), (4, 2, 0)], [(4, 1, 0), (16, 3, 0)],
               [(49, 0, 0), (169, 5, 0)], [(64, 5, 0), (144, 0, 0)], [(169, 3, 0), (361, 1, 0)]],
        M=[2, 2, 2, 2, 2, 2],
        new=[[1, 2, 3, 4], [0, 2, 5], [1, 3], [0, 5], [5, 0], [3, 1]],
        old=[[], [], [], [], [], []],
        stats=[(0, 0)]),
    2: dict(
        pools=[[(1, 1, 0), (9, 2, 0)], [(1, 0, 0), (4, 2, 0)], [(4, 1, 0), (9, 0, 0), (16, 3, 0)],
               [(16, 2, 0), (25, 4, 0), (36, 1, 1)], [(25, 3, 0), (64, 5, 0), (81, 2, 0)],
               [(169, 3, 0), (289, 2, 0), (361, 1, 0)]],
        M=[2, 2, 3, 2, 3, 3],
        new=[[], [], [0, 4, 5], [2, 4], [3, 2], [2]],
        old=[[1, 2], [0, 2, 5], [1, 3], [5], [5], [3, 1]],
        stats=[(0, 0), (13, 8)]),
    3: dict(
        pools=[[(1, 1, 0), (9, 2, 0)], [(1, 0, 0), (4, 2, 0)], [(4, 1, 0), (9, 0, 0), (16, 3, 0)],
               [(16, 2, 0), (25, 4, 0), (36, 1, 0)], [(25, 3, 0), (64, 5, 0), (81, 2, 0)],
               [(64, 4, 0), (169, 3, 0), (289, 2, 0)]],
        M=[2, 2, 3, 3, 3, 3],
        new=[[], [3], [], [1], [], [4]],
        old=[[1, 2], [0, 2], [1, 0, 3, 4, 5], [2, 4, 5], [3, 5, 2], [3, 2]],
        stats=[(0, 0), (13, 8), (17, 1)]),
}

# final graphs (ids by ascending distance) and total join pairs
FINAL = {
    1: ([[1, 2], [0, 2], [1, 3], [0, 5], [5, 0], [3, 1]], 0),
    2: ([[1, 2], [0, 2], [1, 0, 3], [2, 4, 1], [3, 5, 2], [3, 2, 1]], 13),
    3: ([[1, 2], [0, 2], [1, 0, 3], [2, 4, 1], [3, 5, 2], [4, 3, 2]], 30),
    4: ([[1, 2], [0, 2], [1, 0, 3], [2, 4, 1], [3, 5, 2], [4, 3, 2]], 37),
}


def test_kat_distances_are_the_hand_values():
    assert INIT_DIST.tolist() == [[1, 9], [1, 4], [4, 16], [49, 169], [64, 144], [169, 361]]


@pytest.mark.parametrize("iters", [1, 2, 3])
def test_kat_update_state(iters):
    st = O.nndescent_state(X, 3, 3, 2, 8, iters, init=INIT, nthreads=1)
    e = EXPECTED[iters]
    for n in range(6):
        c = int(st["pool_cnt"][n])
        got = [(int(st["pool_dist"][n, j]), int(st["pool_ids"][n, j]), int(st["pool_flag"][n, j]))
               for j in range(c)]
        assert got == e["pools"][n], f"node {n} pool"
    assert st["M"].tolist() == e["M"]
    assert [[int(v) for v in l] for l in st["nn_new"]] == e["new"]
    assert [[int(v) for v in l] for l in st["nn_old"]] == e["old"]
    assert st["stats"] == e["stats"]


def test_kat_reference_sweep_round0():
    """The reference's single-thread sweep (sweep=True): sampling node n ends
    with make_heap of n's pool (src/component_refine.cpp:212), so a later
    node reads a heap leaf as n's pool.back().  A sorted pool of two,
    [a0, a1], heapifies to [a1, a0]: the leaf is the smaller entry, so the
    leaves are 1, 1, 4, 49, 64, 169 (worsts 9, 4, 16, 169, 144, 361).  Round
    0's reverse tests, leaf for neighbours with a smaller id:
      n=1: nb 0 leaf 1: 1 > 1 no;      n=2: nb 1 leaf 1: 4 > 1 YES;
      n=3: nb 0: 49 > 1 YES;           n=4: nb 0: 144 > 1 YES;
      n=5: nb 3 leaf 49: 169 > 49 YES, nb 1: 361 > 1 YES;
    neighbours with a larger id read the worst as before (all no)."""
    st = O.nndescent_state(X, 3, 3, 2, 8, 1, init=INIT, nthreads=1, sweep=True)
    assert [[int(v) for v in l] for l in st["nn_new"]] == [
        [1, 2, 3, 4], [0, 2, 2, 5], [1, 3], [0, 5, 5], [5, 0], [3, 1]]
    assert st["M"].tolist() == [2] * 6


@pytest.mark.parametrize("slots", [0, 64])
@pytest.mark.parametrize("iters", [1, 2, 3, 4])
def test_kat_final_graph(iters, slots):
    # the candidate-slot mode (threshold and pool-id filter from the start of
    # the join, winners merged after it) keeps the same top-L here: the
    # targets never receive two candidates in one slot
    ids, dist, deg, st = O.nndescent(X, 3, 3, 2, 8, iters, init=INIT, cand_slots=slots, nthreads=1,
                                     with_stats=True)
    want, pairs = FINAL[iters]
    assert [list(map(int, ids[n, :deg[n]])) for n in range(6)] == want
    for n in range(6):
        assert np.array_equal(dist[n, :deg[n]], (X[ids[n, :deg[n]], 0] - X[n, 0]) ** 2)
    assert sum(p for p, _ in st) == pairs


def test_kat_slot_collisions_absent():
    """The slot-mode check above is meaningful only if no target sees two
    distinct candidates in one slot of a 64-slot table."""
    def fmix32(h):
        h ^= h >> 16
        h = (h * 0x85EBCA6B) & 0xFFFFFFFF
        h ^= h >> 13
        h = (h * 0xC2B2AE35) & 0xFFFFFFFF
        return h ^ (h >> 16)

    # slot of candidate j in target t's table (the library's cand_slot)
    for t in range(6):
        slots = [fmix32((j * 0x9E3779B1 + t * 0x7F4A7C15) & 0xFFFFFFFF) & 63 for j in range(6) if j != t]
        assert len(set(slots)) == len(slots)


@pytest.mark.gpu
@pytest.mark.parametrize("iters", [1, 2, 3, 4])
def test_kat_gpu_build(ctx, iters):
    """The device build of the hand-worked case (FP32 join, 64 slots)."""
    ids, dist, deg, st = ctx.nndescent_build(X, 3, 3, 2, 8, iters, seed=0, cand_slots=64,
                                             init_ids=INIT_IDS)
    want, pairs = FINAL[iters]
    assert [list(map(int, ids[n, :deg[n]])) for n in range(6)] == want
    assert st["pairs"] == pairs
