"""CPU tests pinning the oracle's wide-BVH collapse (orc_wide_collapse,
restating builder/cuda/wide_gpu_builder.h:39-160).

* known answer on a hand-built tree: the children of each wide node in the
  reference's depth-first order (offset+1 popped before offset+0);
* closed form on built trees: every binary node maps to exactly one place
  in the wide tree -- the root and each inner node at a depth that is a
  multiple of log2(W) become exactly one wide node and (except the root)
  exactly one inner child pointing at it, each leaf exactly one leaf child
  with its bounds, offset and count; every other inner node is absorbed;
  so the wide node count is 1 + the number of inner nodes at those depths,
  and the leaf children's ranges partition the primIDs;
* a planted mistake (depth multiple taken as log2(W)+1) fails the closed
  form.
"""
import numpy as np
import pytest

F32 = np.float32


def _node(lo, hi, off, cnt):
    return (np.asarray(lo, F32), np.asarray(hi, F32), np.uint64(off | (cnt << 48)))


def _hand_tree(oracle_mod):
    # 0 -> (2,3); 2 -> leaves (4,5); 3 -> (6,7); 6 -> leaves (8,9); 7 -> leaves (10,11)
    nodes = np.zeros(12, dtype=oracle_mod.NODE_DTYPE)
    spec = {0: (2, 0), 2: (4, 0), 3: (6, 0), 4: (0, 1), 5: (1, 1), 6: (8, 0), 7: (10, 0),
            8: (2, 1), 9: (3, 1), 10: (4, 2), 11: (6, 1)}
    for i, (off, cnt) in spec.items():
        nodes[i] = _node([i, i, i], [i + 0.5] * 3, off, cnt)
    return oracle_mod.Tree(nodes, np.arange(7, dtype=np.uint32))


def test_wide_collapse_known_answer(oracle_mod):
    t = _hand_tree(oracle_mod)
    w4 = oracle_mod.wide_collapse(t, 4)
    valid, off, cnt = oracle_mod.wide_fields(w4)
    # wide roots: 0 (id 0), then the inner nodes at depth 2: 6 (id 1), 7 (id 2)
    assert w4.shape == (3, 4)
    # wide 0: walk 0 -> push 2,3; pop 3 -> push 6,7; pop 7 (root), pop 6 (root), pop 2 -> push 4,5; pop 5, pop 4
    assert valid[0].tolist() == [True] * 4
    assert [int(w4[0, j]["lo"][0]) for j in range(4)] == [7, 6, 5, 4]
    assert off[0].tolist() == [2, 1, 1, 0] and cnt[0].tolist() == [0, 0, 1, 1]
    # wide 1 (node 6): leaves 9 then 8; wide 2 (node 7): leaves 11 then 10
    assert [int(w4[1, j]["lo"][0]) for j in range(2)] == [9, 8] and valid[1].tolist() == [True, True, False, False]
    assert [int(w4[2, j]["lo"][0]) for j in range(2)] == [11, 10]
    assert np.all(np.isinf(w4[1, 2]["lo"])) and cnt[1, 2] == 0xFFFF and off[1, 2] == 0xFFFFFFFF
    # W = 8: everything within three levels of the root is one node
    w8 = oracle_mod.wide_collapse(t, 8)
    v8, o8, c8 = oracle_mod.wide_fields(w8)
    assert w8.shape == (1, 8)
    assert [int(w8[0, j]["lo"][0]) for j in range(6)] == [11, 10, 9, 8, 5, 4]
    assert v8[0].tolist() == [True] * 6 + [False] * 2


def _depths(tree):
    d = np.full(len(tree.nodes), -1, np.int64)
    d[0] = 0
    stack = [0]
    while stack:
        i = stack.pop()
        if tree.counts[i] == 0:
            for c in (0, 1):
                d[tree.offsets[i] + c] = d[i] + 1
                stack.append(tree.offsets[i] + c)
    return d


def check_collapse(oracle_mod, tree, wide, W, logw=None):
    """The closed-form mapping; returns nothing, asserts."""
    logw = logw or (2 if W == 4 else 3)
    d = _depths(tree)
    reach = d >= 0
    inner = (tree.counts == 0) & reach
    roots = np.nonzero(((d % logw == 0) & inner & (d > 0)))[0]
    assert wide.shape[0] == 1 + len(roots)
    valid, off, cnt = oracle_mod.wide_fields(wide)
    assert np.all(valid.sum(1) <= W) and np.all(valid.sum(1) >= 1)
    # every valid child matches exactly one binary node (leaf or wide root)
    leaf_ids = np.nonzero((tree.counts > 0) & reach)[0]
    leaf_key = {(int(tree.offsets[i]), int(tree.counts[i])): i for i in leaf_ids}
    seen_leaf, seen_inner = set(), set()
    root_of = {0: 0}
    root_order = [0] + sorted(int(r) for r in roots)
    for wid, r in enumerate(root_order):
        root_of[r] = wid
    for wid in range(wide.shape[0]):
        for j in range(W):
            if not valid[wid, j]:
                assert np.all(np.isinf(wide[wid, j]["lo"])) and np.all(wide[wid, j]["lo"] > 0)
                continue
            if cnt[wid, j] > 0:
                i = leaf_key[(int(off[wid, j]), int(cnt[wid, j]))]
                assert i not in seen_leaf
                seen_leaf.add(i)
            else:
                i = root_order[int(off[wid, j])]
                assert i != 0 and i not in seen_inner
                seen_inner.add(i)
            assert np.array_equal(wide[wid, j]["lo"], tree.nodes[i]["lo"])
            assert np.array_equal(wide[wid, j]["hi"], tree.nodes[i]["hi"])
    assert seen_leaf == set(int(i) for i in leaf_ids)
    assert seen_inner == set(int(r) for r in roots)
    # the leaf ranges partition the primIDs
    covered = np.zeros(len(tree.prim_ids), np.int64)
    for (o, c) in leaf_key:
        covered[o:o + c] += 1
    assert np.all(covered == 1)


@pytest.mark.parametrize("builder,W", [("sm", 4), ("sm", 8), ("radix", 4), ("sah", 8), ("cpu1", 4)])
def test_wide_collapse_closed_form(oracle_mod, builder, W):
    P = oracle_mod.lcg_uniform(9, 20_000)
    B = oracle_mod.points_to_boxes(P)
    tree = {"sm": lambda: oracle_mod.sm_build(B, 8), "radix": lambda: oracle_mod.radix_build(B, 10, 4),
            "sah": lambda: oracle_mod.sah_build(B, 2, 16), "cpu1": lambda: oracle_mod.sm_build(B, 1)}[builder]()
    check_collapse(oracle_mod, tree, oracle_mod.wide_collapse(tree, W), W)


def test_wide_collapse_single_leaf(oracle_mod):
    B = oracle_mod.points_to_boxes(oracle_mod.lcg_uniform(1, 5))
    tree = oracle_mod.sm_build(B, 8)
    assert len(tree.nodes) == 1
    w = oracle_mod.wide_collapse(tree, 4)
    valid, off, cnt = oracle_mod.wide_fields(w)
    assert w.shape == (1, 4) and valid[0].tolist() == [True, False, False, False]
    assert off[0, 0] == 0 and cnt[0, 0] == 5


def test_wide_closed_form_detects_wrong_depth_rule(oracle_mod):
    P = oracle_mod.lcg_uniform(10, 5_000)
    tree = oracle_mod.sm_build(oracle_mod.points_to_boxes(P), 4)
    with pytest.raises((AssertionError, KeyError, IndexError)):
        check_collapse(oracle_mod, tree, oracle_mod.wide_collapse(tree, 4), 4, logw=3)
