"""CPU tests pinning the oracle's fixed-box query (orc_box_query, restating
traversal/fixedBoxQuery.h).

* sets: for every query box, the reported prims are exactly those whose box
  overlaps it (numpy brute force on closed boxes), and counts match;
* order, closed form on a Morton tree (whose leaves lie in primID-array
  order along a first-child-first walk): the binary query reports the
  overlapping prims in primID-array order; the wide query (children of each
  wide node in the collapse's second-child-first order) reports its leaves
  in reverse array order, each leaf's prims in array order;
* truncation: with max_out smaller than a count, the first max_out entries
  of the full list are kept and the count is the full count.
"""
import numpy as np
import pytest


def _overlap(pb, q):
    return ~np.any((pb[:, 3:] < q[:3]) | (q[3:] < pb[:, :3]), axis=1)


def _queries(rng, n, size):
    c = rng.random((n, 3)).astype(np.float32)
    h = (rng.random((n, 3)) * size).astype(np.float32)
    return np.concatenate([c - h, c + h], axis=1).astype(np.float32)


@pytest.fixture(scope="module")
def setup(oracle_mod):
    rng = np.random.default_rng(3)
    P = oracle_mod.lcg_uniform(12, 20_000)
    # small boxes around the points, so prim boxes overlap each other
    B = np.concatenate([P - 0.002, P + 0.002], axis=1).astype(np.float32)
    Q = _queries(rng, 300, 0.05)
    tree = oracle_mod.radix_build(B, 10, 4)
    return P, B, Q, tree


def test_box_query_sets_and_binary_order(oracle_mod, setup):
    P, B, Q, tree = setup
    ids, cnt = oracle_mod.box_query(tree, B, Q, 4096)
    pos = np.empty(len(tree.prim_ids), np.int64)
    pos[tree.prim_ids] = np.arange(len(tree.prim_ids))
    for q in range(len(Q)):
        want = np.nonzero(_overlap(B, Q[q]))[0]
        assert cnt[q] == len(want)
        got = ids[q, :cnt[q]]
        assert np.array_equal(np.sort(got), want)
        assert np.array_equal(got, want[np.argsort(pos[want])])   # primID-array order


@pytest.mark.parametrize("W", [4, 8])
def test_box_query_wide_order(oracle_mod, setup, W):
    P, B, Q, tree = setup
    wide = oracle_mod.wide_collapse(tree, W)
    ids, cnt = oracle_mod.box_query(tree, B, Q, 4096, wide=wide, width=W)
    pos = np.empty(len(tree.prim_ids), np.int64)
    pos[tree.prim_ids] = np.arange(len(tree.prim_ids))
    # leaf (offset) of each array position
    leaf_start = np.zeros(len(tree.prim_ids), np.int64)
    for off, c in zip(tree.offsets[tree.counts > 0], tree.counts[tree.counts > 0]):
        leaf_start[off:off + c] = off
    for q in range(len(Q)):
        want = np.nonzero(_overlap(B, Q[q]))[0]
        assert cnt[q] == len(want)
        p = pos[want]
        order = np.lexsort((p, -leaf_start[p]))   # leaves descending, inside a leaf ascending
        assert np.array_equal(ids[q, :cnt[q]], want[order])


def test_box_query_truncation(oracle_mod, setup):
    P, B, Q, tree = setup
    full, cnt = oracle_mod.box_query(tree, B, Q, 4096)
    few, cnt2 = oracle_mod.box_query(tree, B, Q, 3)
    assert np.array_equal(cnt, cnt2)
    for q in range(len(Q)):
        k = min(3, int(cnt[q]))
        assert np.array_equal(few[q, :k], full[q, :k])
