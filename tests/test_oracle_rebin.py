"""CPU tests pinning the oracle's robust rebinning Morton builder.

The reference builder is CUDA-only, so orc_rebin_build is pinned by:

* an independent pure-Python restatement (float32 numpy scalars for the
  quantizer, written from rebinMortonBuilder.h, not from the C++ oracle) on
  small clustered inputs with outliers, which must give the same primID
  order and the same tree;
* reduction to a known case: with no range of equal keys larger than the
  leaf threshold no rebinning happens, and the tree equals the 30-bit radix
  tree (orc_radix_hierarchy, itself pinned by the Morton/sort fixtures);
* closed forms: identical points are keyed by local index, so the root of
  100 splits 64/36 and the prim order stays the input order; a tight cluster
  beside a far outlier is re-keyed in its own domain in the thread phase;
* TreeChecker invariants and kNN equal to brute force on a Gaussian mixture
  with outliers (the C4 generator).
"""
import numpy as np
import pytest

F32 = np.float32


# ---------------------------------------------------------------- pure Python
def _spread(v):
    r = 0
    for i in range(10):
        r |= ((v >> i) & 1) << (3 * i)
    return r


def _quant(lo, hi):
    ext = max(max(F32(hi[0] - lo[0]), F32(hi[1] - lo[1])), F32(hi[2] - lo[2]))
    return lo, F32(1024.0) * (F32(1.0) / max(ext, F32(1e-20)))


def _key(c, q):
    lo, scale = q
    cell = [min(int(F32((c[d] - lo[d])) * scale), 1023) for d in range(3)]
    return _spread(cell[0]) | (_spread(cell[1]) << 1) | (_spread(cell[2]) << 2)


def _center(b):
    return [F32(b[d] + b[d + 3]) / F32(2) for d in range(3)]


def _domain(boxes, ids):
    cs = [_center(boxes[i]) for i in ids]
    lo = [min(c[d] for c in cs) for d in range(3)]
    hi = [max(c[d] for c in cs) for d in range(3)]
    return lo, hi, cs


def _find_split(keys, b, e):
    fk, lk = keys[b], keys[e - 1]
    if fk == lk:
        return None
    nm = 32 - (fk ^ lk).bit_length()
    search = lk & ((0xFFFFFFFF << (31 - nm)) & 0xFFFFFFFF)
    lo, hi = b, e
    while hi > lo:
        mid = (lo + hi) // 2
        if keys[mid] < search:
            lo = mid + 1
        else:
            hi = mid
    return lo


def py_rebin(boxes, T):
    """Returns (primIDs, canonical nested tuple) of the rebin builder."""
    boxes = [list(map(F32, b)) for b in boxes]
    valid = [i for i, b in enumerate(boxes) if not (b[0] > b[3])]
    lo, hi, cs = _domain(boxes, valid)
    q = _quant(lo, hi)
    pairs = sorted(((_key(_center(boxes[i]), q), i) for i in valid), key=lambda p: p[0])
    keys = [k for k, _ in pairs]
    ids = [i for _, i in pairs]

    def rec(b, e):
        n = e - b
        if n <= T:
            return tuple(sorted(ids[b:e]))
        if n > 32:
            while True:
                s = _find_split(keys, b, e)
                if s is not None:
                    break
                dlo, dhi, dcs = _domain(boxes, ids[b:e])
                deg = dlo == dhi
                dq = _quant(dlo, dhi)
                nk = [(i if deg else _key(dcs[i], dq), ids[b + i]) for i in range(n)]
                if len(set(k for k, _ in nk)) == 1:
                    nk = [(i, v) for i, (_, v) in enumerate(nk)]
                nk.sort(key=lambda p: p[0])   # stable
                keys[b:e] = [k for k, _ in nk]
                ids[b:e] = [v for _, v in nk]
        else:
            s = _find_split(keys, b, e)
            if s is None:
                dlo, dhi, dcs = _domain(boxes, ids[b:e])
                deg = dlo == dhi
                dq = _quant(dlo, dhi)
                for i in range(n):
                    keys[b + i] = i if deg else _key(dcs[i], dq)
                for i in range(b, e):          # bubbleSort (rebinMortonBuilder.h:758-770)
                    for j in range(i + 1, e):
                        if keys[j] < keys[i]:
                            keys[i], keys[j] = keys[j], keys[i]
                            ids[i], ids[j] = ids[j], ids[i]
                s = _find_split(keys, b, e)
                if s is None:
                    return tuple(sorted(ids[b:e]))
        return ("node", rec(b, s), rec(s, e))

    root = rec(0, len(ids)) if ids else None
    return ids, root


def _tree_tuple(tree):
    def rec(i):
        cnt, off = int(tree.counts[i]), int(tree.offsets[i])
        if cnt:
            return tuple(sorted(tree.prim_ids[off:off + cnt].tolist()))
        return ("node", rec(off), rec(off + 1))
    return rec(0)


def _clustered(n_per, centres, sigma, seed, outliers=()):
    rng = np.random.default_rng(seed)
    pts = [np.asarray(c, dtype=np.float64) + sigma * rng.standard_normal((n_per, 3)) for c in centres]
    pts.append(np.asarray(outliers, dtype=np.float64).reshape(-1, 3))
    p = np.concatenate(pts).astype(F32)
    return p[rng.permutation(len(p))]


@pytest.mark.parametrize("T,seed", [(1, 0), (4, 1), (8, 2)])
def test_rebin_matches_pure_python(oracle_mod, T, seed):
    P = _clustered(80, [[0, 0, 0], [5, 5, 5], [5, 0, 5]], 1e-3, seed,
                   outliers=[[1e6, 0, 0], [-1e6, 1e6, 0]])
    # a few exact duplicates force the degenerate (local index) keys
    P = np.concatenate([P, np.tile(P[:1], (40, 1)), np.tile(P[5:6], (7, 1))])
    B = oracle_mod.points_to_boxes(P)
    tree = oracle_mod.rebin_build(B, T)
    ids, root = py_rebin(B, T)
    assert tree.prim_ids.tolist() == ids
    assert _tree_tuple(tree) == root
    assert oracle_mod.check_tree(tree, B)[0] == 0


def test_rebin_equals_radix30_without_equal_ranges(oracle_mod):
    P = oracle_mod.lcg_uniform(3, 50_000)
    B = oracle_mod.points_to_boxes(P)
    keys, ids = oracle_mod.morton_keys(B, 10)
    keys, ids = oracle_mod.sort_pairs(keys, ids)
    # the largest run of equal keys is far below the leaf threshold
    _, runs = np.unique(keys, return_counts=True)
    assert runs.max() <= 8
    a = oracle_mod.rebin_build(B, 8)
    b = oracle_mod.radix_build(B, 10, 8)
    assert np.array_equal(a.prim_ids, b.prim_ids)
    assert np.array_equal(oracle_mod.canonical(a), oracle_mod.canonical(b))


def test_rebin_identical_points(oracle_mod):
    B = np.tile(np.array([[1, 2, 3, 1, 2, 3]], dtype=F32), (100, 1))
    tree = oracle_mod.rebin_build(B, 1)
    assert tree.prim_ids.tolist() == list(range(100))
    off = int(tree.offsets[0])
    sizes = sorted(len(_flat(tree, off + j)) for j in range(2))
    assert sizes == [36, 64]
    code, depth, leaves = oracle_mod.check_tree(tree, B, 1)
    assert code == 0 and leaves == 100


def _flat(tree, i):
    cnt, off = int(tree.counts[i]), int(tree.offsets[i])
    if cnt:
        return tree.prim_ids[off:off + cnt].tolist()
    return _flat(tree, off) + _flat(tree, off + 1)


def test_rebin_thread_phase_cluster_beside_outlier(oracle_mod):
    rng = np.random.default_rng(11)
    P = np.concatenate([rng.random((10, 3)) * 1e-3, [[1e6, 1e6, 1e6]]]).astype(F32)
    B = oracle_mod.points_to_boxes(P)
    tree = oracle_mod.rebin_build(B, 1)
    ids, root = py_rebin(B, 1)
    assert tree.prim_ids.tolist() == ids and _tree_tuple(tree) == root
    # the cluster is split spatially: every leaf holds one point
    assert oracle_mod.check_tree(tree, B, 1)[0] == 0


def test_rebin_gaussian_mixture_queries(oracle_mod):
    from cubql_amd import synth
    P = synth.gaussian_mixture(20_000, seed=50)
    Q = synth.gaussian_mixture(300, seed=51)
    B = oracle_mod.points_to_boxes(P)
    tree = oracle_mod.rebin_build(B, 8)
    code, depth, leaves = oracle_mod.check_tree(tree, B, 32)
    assert code == 0
    ids, d2, _ = oracle_mod.knn(tree, P, Q, 16)
    bi, bd = oracle_mod.knn_brute(P, Q, 16)
    assert np.array_equal(ids, bi) and np.array_equal(d2, bd)


# ------------------------------------------------------- reduction to radix30
def isolated_cluster(n, seed, centre=(5.0, 5.0, 5.0), extent=1e-3):
    """n points uniform in a cube of side `extent` around `centre`, far
    outliers at +-1e6 (so the global 30-bit quantizer puts the whole cluster
    in one cell) and the permutation that interleaves them in the input.
    Returns (points, cluster_ids): cluster_ids are the cluster's primIDs in
    ascending order."""
    rng = np.random.default_rng(seed)
    c = np.asarray(centre) + extent * (rng.random((n, 3)) - 0.5)
    far = np.array([[1e6, 1e6, 1e6], [-1e6, -1e6, -1e6], [1e6, -1e6, 3e5],
                    [-7e5, 2e5, 1e6]])
    P = np.concatenate([c, far]).astype(F32)
    perm = rng.permutation(len(P))
    P = P[perm]
    cluster_ids = np.sort(np.nonzero(perm < n)[0]).astype(np.uint32)
    return P, cluster_ids


def find_subtree(tree, prim_set):
    """Index of the node whose prims are exactly `prim_set`, or -1."""
    want = frozenset(int(x) for x in prim_set)

    def rec(i):
        cnt, off = int(tree.counts[i]), int(tree.offsets[i])
        if cnt:
            s = frozenset(tree.prim_ids[off:off + cnt].tolist())
            return s, (i if s == want else -1)
        a, fa = rec(off)
        b, fb = rec(off + 1)
        s = a | b
        return s, (i if s == want else max(fa, fb))
    return rec(0)[1]


def subtree_tuple(tree, i, remap):
    cnt, off = int(tree.counts[i]), int(tree.offsets[i])
    if cnt:
        return tuple(sorted(remap[int(p)] for p in tree.prim_ids[off:off + cnt]))
    return ("node", subtree_tuple(tree, off, remap), subtree_tuple(tree, off + 1, remap))


def subtree_prim_order(tree, i, remap):
    """primIDs of the subtree as stored (a Morton subtree covers one
    contiguous primID range), mapped through `remap`."""
    def span(j):
        cnt, off = int(tree.counts[j]), int(tree.offsets[j])
        if cnt:
            return off, off + cnt
        a, b = span(off), span(off + 1)
        return min(a[0], b[0]), max(a[1], b[1])
    b, e = span(i)
    return [remap[int(p)] for p in tree.prim_ids[b:e]]


def check_cluster_reduction(oracle_mod, build, P, cluster_ids, T):
    """The rebin tree's cluster subtree equals radix30 over the cluster alone
    (rebinMortonBuilder.h:1305-1405 broad phase, :772-886 thread phase: the
    range is re-keyed with the quantizer of its own centroid bounds, the same
    formula radix.h:41-70 applies to a whole input, then re-sorted; with
    distinct keys in that domain no further rebin happens and every sort
    order is the key order)."""
    B = oracle_mod.points_to_boxes(P)
    C = P[cluster_ids]
    CB = oracle_mod.points_to_boxes(C)
    # preconditions: one global cell holds exactly the cluster, and the
    # cluster's keys in its own domain are distinct
    gk, gi = oracle_mod.morton_keys(B, 10)
    ck = gk[np.isin(gi, cluster_ids)]
    assert len(np.unique(ck)) == 1 and np.sum(gk == ck[0]) == len(cluster_ids)
    lk, _ = oracle_mod.morton_keys(CB, 10)
    assert len(np.unique(lk)) == len(cluster_ids)
    tree = build(B, T)
    node = find_subtree(tree, cluster_ids)
    assert node >= 0, "no node holds exactly the cluster"
    remap = {int(g): j for j, g in enumerate(cluster_ids)}
    alone = oracle_mod.radix_build(CB, 10, T)
    ident = {j: j for j in range(len(cluster_ids))}
    assert subtree_prim_order(tree, node, remap) == alone.prim_ids.tolist()
    assert subtree_tuple(tree, node, remap) == subtree_tuple(alone, 0, ident)
    return tree


@pytest.mark.parametrize("n,T,seed", [(2000, 1, 0), (2000, 8, 1), (600, 4, 2),   # broad phase
                                      (33, 1, 3), (32, 1, 4), (20, 2, 5), (9, 1, 6)])  # thread phase
def test_rebin_cluster_reduces_to_radix30(oracle_mod, n, T, seed):
    P, cl = isolated_cluster(n, seed)
    check_cluster_reduction(oracle_mod, oracle_mod.rebin_build, P, cl, T)


def test_rebin_cluster_reduction_detects_parent_domain(oracle_mod):
    """A restatement that re-keyed the cluster in the quantizer of the whole
    input (instead of the cluster's own centroid bounds) would leave every key
    equal; the reduction check must then fail.  Here: the radix30 tree of
    the whole input, whose cluster range is split in the middle, fails it."""
    P, cl = isolated_cluster(300, 7)
    with pytest.raises(AssertionError):
        check_cluster_reduction(oracle_mod, lambda b, t: oracle_mod.radix_build(b, 10, t), P, cl, 1)
