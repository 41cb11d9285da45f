"""CPU tests pinning the oracle's binned SAH builder (orc_sah_build).

The reference's SAH builder is CUDA-only (no host path to compile), so its
restatement is pinned by what the mathematics fixes:

* brute-force optimality: at every node of an oracle tree, an independent
  numpy evaluation that bins the node's prims, forms every split's two sides
  directly from the prim boxes (not from bin prefix unions) and takes the
  first minimum over [leaf, dim 0..2 x bin 0..numBins-1] predicts exactly the
  node's children (or that it is a leaf);
* closed forms: two separated clusters split apart at the root; identical
  boxes give one leaf (no split has two non-empty sides) or, under a leaf
  cap, middle splits into leaves of at most the cap;
* TreeChecker invariants and leaf caps on random boxes and triangle boxes;
* query results over SAH trees equal the brute-force closest point.
"""
import numpy as np
import pytest

F32 = np.float32


def _area(lo, hi):
    sx, sy, sz = hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]
    return sx * sy + sx * sz + sy * sz


def _bins(B, nb):
    c = (B[:, :3] + B[:, 3:]) / F32(2)
    lo, hi = c.min(0), c.max(0)
    out = []
    for d in range(3):
        if hi[d] > lo[d]:
            pd = F32(0.5) * (B[:, d] + B[:, 3 + d])
            f = (pd - lo[d]) / (hi[d] - lo[d] + F32(1e-20)) * F32(nb)
            b = np.where(f >= nb - 1, nb - 1, np.where(f > 0, np.floor(f), 0)).astype(np.int64)
        else:
            b = np.zeros(len(B), dtype=np.int64)
        out.append(b)
    return out


def np_decide(B, nb, allow_leaf=True):
    """First minimum of the binned SAH cost sequence, computed from prim sets."""
    B = B.astype(F32)
    n = len(B)
    best, choice = F32(np.inf), (-1, 0)
    leaf_cost = _area(B[:, :3].min(0), B[:, 3:].max(0)) * F32(n)
    if allow_leaf and leaf_cost < best:
        best = leaf_cost
    bins = _bins(B, nb)
    for d in range(3):
        for b in range(nb):
            L = bins[d] < b
            nl = int(L.sum())
            if nl == 0 or nl == n:
                continue
            R = ~L
            cost = (_area(B[L, :3].min(0), B[L, 3:].max(0)) * F32(nl)
                    + _area(B[R, :3].min(0), B[R, 3:].max(0)) * F32(n - nl))
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


def _random_boxes(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3), dtype=np.float32)
    h = (rng.random((n, 3), dtype=np.float32) * F32(0.02)).astype(F32)
    return np.concatenate([c - h, c + h], axis=1).astype(F32)


@pytest.mark.parametrize("nb,T", [(16, 1), (32, 4), (8, 2)])
def test_sah_every_node_is_the_first_minimum(oracle_mod, nb, T):
    B = _random_boxes(1500, seed=nb + T)
    tree = oracle_mod.sah_build(B, T, nb, 1 << 15)
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
        (d, b), bins = np_decide(B[prims], nb)
        if cnt:
            assert d == -1, f"node {i}: leaf of {cnt} but a split is cheaper"
            continue
        assert d >= 0, f"node {i}: split but the leaf is not beaten"
        left = set(prims[bins[d] < b].tolist())
        l0 = set(_subtree_prims(tree, off).tolist())
        l1 = set(_subtree_prims(tree, off + 1).tolist())
        assert l0 == left and l1 == set(prims.tolist()) - left
        checked += 1
        st += [off, off + 1]
    assert checked > 100


def test_sah_root_decision_matches_numpy(oracle_mod):
    for seed in range(20):
        B = _random_boxes(300, seed=100 + seed)
        (d, b), _ = np_decide(B, 16)
        d0, b0 = oracle_mod.sah_decide(B, 16)
        assert d0 == d and (d < 0 or b0 == b)


def test_sah_two_clusters_split_at_root(oracle_mod):
    a = _random_boxes(200, seed=1)
    b = _random_boxes(300, seed=2)
    b[:, [0, 3]] += F32(100.0)
    B = np.concatenate([a, b])
    tree = oracle_mod.sah_build(B, 1, 16)
    l0 = set(_subtree_prims(tree, int(tree.offsets[0])).tolist())
    l1 = set(_subtree_prims(tree, int(tree.offsets[0]) + 1).tolist())
    assert {frozenset(l0), frozenset(l1)} == {frozenset(range(200)), frozenset(range(200, 500))}


def test_sah_identical_boxes(oracle_mod):
    B = np.tile(np.array([[0, 0, 0, 1, 1, 1]], dtype=F32), (100, 1))
    # no split has two non-empty sides: one leaf (reference behaviour)
    tree = oracle_mod.sah_build(B, 1, 16, 1 << 15)
    assert len(tree.nodes) == 1 and tree.counts[0] == 100
    # leaf cap: middle splits down to leaves of <= 4, balanced
    tree = oracle_mod.sah_build(B, 1, 16, 4)
    code, depth, leaves = oracle_mod.check_tree(tree, B, 4)
    assert code == 0 and depth == 6 and leaves == 32
    assert oracle_mod.sah_decide(B, 16, allow_leaf=False)[0] == -2


def test_sah_leaf_cap_random(oracle_mod):
    # 60 copies of one box can never be split by a plane: a leaf of >= 60
    B = np.concatenate([_random_boxes(3000, seed=7),
                        np.tile(_random_boxes(1, seed=8), (60, 1))])
    free = oracle_mod.sah_build(B, 1, 16, 1 << 15)
    capped = oracle_mod.sah_build(B, 4, 32, 4)
    assert oracle_mod.check_tree(free, B)[0] == 0
    assert oracle_mod.check_tree(capped, B, 4)[0] == 0
    assert free.counts.max() >= 60 and capped.counts.max() <= 4


def test_sah_invalid_boxes_skipped(oracle_mod):
    B = _random_boxes(500, seed=3)
    B[::7, 0], B[::7, 3] = F32(1.0), F32(0.0)   # lower.x > upper.x: invalid
    tree = oracle_mod.sah_build(B, 2, 16)
    assert oracle_mod.check_tree(tree, B)[0] == 0
    assert len(tree.prim_ids) == 500 - len(range(0, 500, 7))


def test_sah_triangle_mesh_and_queries(oracle_mod):
    from cubql_amd import synth
    V, F = synth.noisy_icosphere(3, seed=4)
    B = oracle_mod.tri_boxes(F, V)
    tree = oracle_mod.sah_build(B, 4, 32, 4)
    assert oracle_mod.check_tree(tree, B, 4)[0] == 0
    Q = synth.mesh_surface_queries(V, F, 300, seed=5)
    ids, d2, _ = oracle_mod.fcp_tris(tree, F, V, Q)
    _, d64 = oracle_mod.fcp_tris_brute_f64(F, V, Q)
    assert np.allclose(np.sqrt(d2), np.sqrt(d64), rtol=1e-5, atol=1e-6)


def test_sah_points_fcp_equals_brute_force(oracle_mod):
    P = oracle_mod.lcg_uniform(0, 4000)
    Q = oracle_mod.lcg_uniform(1, 500, -0.1, 1.1)
    tree = oracle_mod.sah_build(oracle_mod.points_to_boxes(P), 8, 16)
    ids, d2 = oracle_mod.fcp(tree, P, Q)
    bi, bd = oracle_mod.fcp_brute(P, Q)
    assert np.array_equal(ids, bi) and np.array_equal(d2, bd)
