"""CPU tests pinning the oracle's edge-length-heuristic builder (orc_elh_build).

The reference's ELH builder (builder/cuda/elh_builder.h) is CUDA-only, so its
restatement is pinned the way the SAH restatement is:

* brute-force optimality: at every node of an oracle tree, an independent
  numpy evaluation that bins the node's prims into 8 bins per axis with
  rel = (c-lo)/(hi-lo), forms every split's two sides directly from the prim
  boxes and takes the first minimum over [leaf, dim 0..2 x bin 0..7] of the
  cost (sum of box extents) x count predicts the node's children;
* the measure is the edge-length sum, not the surface area: on boxes where
  the two heuristics disagree the ELH decision is the edge-length one;
* no epsilon in the bin denominator: a node whose centroids span 1e-25 is
  still binned and split (the SAH formula's +1e-20 would put every prim in
  bin 0 and make a leaf);
* TreeChecker invariants, leaf caps and brute-force closest points.
"""
import numpy as np
import pytest

F32 = np.float32


def _edges(lo, hi):
    s = F32(0)
    for d in range(3):
        s = F32(s + F32(hi[d] - lo[d]))
    return s


def _area(lo, hi):
    sx, sy, sz = hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]
    return sx * sy + sx * sz + sy * sz


def _bins(B, nb, eps):
    c = (B[:, :3] + B[:, 3:]) / F32(2)
    lo, hi = c.min(0), c.max(0)
    out = []
    for d in range(3):
        if hi[d] > lo[d]:
            pd = F32(0.5) * (B[:, d] + B[:, 3 + d])
            f = (pd - lo[d]) / (F32(hi[d] - lo[d]) + F32(eps)) * F32(nb)
            b = np.where(f >= nb - 1, nb - 1, np.where(f > 0, np.floor(f), 0)).astype(np.int64)
        else:
            b = np.zeros(len(B), dtype=np.int64)
        out.append(b)
    return out


def np_decide(B, measure=_edges, nb=8, eps=0.0, allow_leaf=True):
    """First minimum of the binned cost sequence, computed from prim sets."""
    B = B.astype(F32)
    n = len(B)
    best, choice = F32(np.inf), (-1, 0)
    leaf_cost = measure(B[:, :3].min(0), B[:, 3:].max(0)) * F32(n)
    if allow_leaf and leaf_cost < best:
        best = leaf_cost
    bins = _bins(B, nb, eps)
    for d in range(3):
        for b in range(nb):
            L = bins[d] < b
            nl = int(L.sum())
            if nl == 0 or nl == n:
                continue
            R = ~L
            cost = (measure(B[L, :3].min(0), B[L, 3:].max(0)) * F32(nl)
                    + measure(B[R, :3].min(0), B[R, 3:].max(0)) * F32(n - nl))
            if cost < best:
                best, choice = cost, (d, b)
    return choice, bins


def _subtree_prims(tree, node):
    out, st = [], [node]
    while st:
        i = st.pop()
        cnt, off = int(tree.counts[i]), int(tree.offsets[i])
        if cnt:
            out.extend(tree.prim_ids[off:off + cnt].tolist())
        else:
            st += [off, off + 1]
    return np.array(sorted(out), dtype=np.int64)


def _random_boxes(n, seed, h=0.02):
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3), dtype=np.float32)
    hh = (rng.random((n, 3), dtype=np.float32) * F32(h)).astype(F32)
    return np.concatenate([c - hh, c + hh], axis=1).astype(F32)


@pytest.mark.parametrize("T,h", [(1, 0.02), (4, 0.0), (8, 0.1)])
def test_elh_every_node_is_the_first_minimum(oracle_mod, T, h):
    B = _random_boxes(1500, seed=T + 11, h=h)
    tree = oracle_mod.elh_build(B, T)
    assert oracle_mod.check_tree(tree, B)[0] == 0
    checked = 0
    st = [0]
    while st:
        i = st.pop()
        prims = _subtree_prims(tree, i)
        cnt, off = int(tree.counts[i]), int(tree.offsets[i])
        if len(prims) <= T:
            assert cnt == len(prims)
            continue
        (d, b), bins = np_decide(B[prims])
        if cnt:
            assert d == -1, f"node {i}: leaf of {cnt} but a split is cheaper"
            continue
        assert d >= 0, f"node {i}: split but the leaf is not beaten"
        left = set(prims[bins[d] < b].tolist())
        assert set(_subtree_prims(tree, off).tolist()) == left
        assert set(_subtree_prims(tree, off + 1).tolist()) == set(prims.tolist()) - left
        checked += 1
        st += [off, off + 1]
    assert checked > 100


def test_elh_measure_is_edge_lengths(oracle_mod):
    """Random sets where the area and edge-length sequences pick different
    planes: the oracle follows edge lengths every time."""
    differ = 0
    for seed in range(60):
        rng = np.random.default_rng(500 + seed)
        B = _random_boxes(40, seed=500 + seed, h=0.3)
        B[:, [1, 4]] *= F32(rng.random() * 4)   # stretch one axis
        e = np_decide(B)[0]
        a = np_decide(B, measure=_area)[0]
        d0, b0 = oracle_mod.elh_decide(B)
        assert d0 == e[0] and (e[0] < 0 or b0 == e[1])
        differ += e != a
    assert differ >= 5


def test_elh_bins_without_epsilon(oracle_mod):
    """Centroids spanning 1e-25: ELH's (hi-lo) denominator bins them and
    splits; SAH's (hi-lo+1e-20) puts all in bin 0 and keeps one leaf."""
    x = F32(1.0) + (np.arange(16, dtype=F32) * F32(2.0 ** -23))  # distinct floats near 1
    B = np.zeros((16, 6), F32)
    B[:, 0] = B[:, 3] = x * F32(1e-25)
    tree = oracle_mod.elh_build(B, 1)
    assert oracle_mod.check_tree(tree, B)[0] == 0 and tree.counts[0] == 0
    assert np_decide(B)[0][0] == 0
    sah = oracle_mod.sah_build(B, 1, 8)
    assert len(sah.nodes) == 1 and sah.counts[0] == 16


def test_elh_identical_boxes_and_cap(oracle_mod):
    B = np.tile(np.array([[0, 0, 0, 1, 1, 1]], dtype=F32), (100, 1))
    tree = oracle_mod.elh_build(B, 1, 1 << 15)
    assert len(tree.nodes) == 1 and tree.counts[0] == 100
    tree = oracle_mod.elh_build(B, 1, 4)
    code, depth, leaves = oracle_mod.check_tree(tree, B, 4)
    assert code == 0 and depth == 6 and leaves == 32


def test_elh_invalid_boxes_and_points(oracle_mod):
    B = _random_boxes(500, seed=3)
    B[::7, 0], B[::7, 3] = F32(1.0), F32(0.0)
    tree = oracle_mod.elh_build(B, 2)
    assert oracle_mod.check_tree(tree, B)[0] == 0
    assert len(tree.prim_ids) == 500 - len(range(0, 500, 7))
    P = oracle_mod.lcg_uniform(0, 4000)
    Q = oracle_mod.lcg_uniform(1, 500, -0.1, 1.1)
    tree = oracle_mod.elh_build(oracle_mod.points_to_boxes(P), 8)
    ids, d2 = oracle_mod.fcp(tree, P, Q)
    bi, bd = oracle_mod.fcp_brute(P, Q)
    assert np.array_equal(ids, bi) and np.array_equal(d2, bd)
