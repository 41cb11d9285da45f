"""CPU tests pinning the oracle (oracle/oracle.cpp) before it is trusted.

What pins each oracle function:

* orc_lcg_uniform      -- golden vectors written by the reference's own
                          UniformPointGenerator (tests/golden/lcg_uniform.npz)
* orc_sm_build/refit   -- golden cpu::spatialMedian trees, node for node and
                          primID for primID (tests/golden/sm_tree_*.npz), plus
                          the C1 tree digest
* orc_fcp / orc_knn    -- golden reference host findClosest / findKNN results
  (mode 0)
* orc_fcp / orc_knn    -- brute force over all points (same float formula,
  (mode 1)                lowest-id ties): exact equality
* orc_fcp_brute        -- FP64 brute force: same ids up to float ties, and the
                          float distance within 2 ulp of the exact one
* orc_check_tree       -- planted mistakes (a moved prim, a shrunk box, a
                          duplicated prim) are each caught
* orc_canonical        -- invariant under child swaps and leaf reordering;
                          changes when a prim moves between leaves
* orc_morton_keys      -- closed form: an independent numpy bit-interleave on
                          hand-picked cells, and the quantizer's edge cells
* orc_radix_hierarchy  -- every inner node splits at the first index whose key
                          has the highest differing bit of its range set
                          (radix.h:467-481); equal-key ranges split in the
                          middle; leaves only at size <= T
* orc_atomicbox_*      -- order-preserving bijection on edge values
"""
import hashlib

import numpy as np
import pytest

from conftest import golden

CASES = ["uniform8k", "dupes", "identical", "adjacent", "boxes_invalid", "clustered"]


def test_lcg_matches_reference_generator(oracle_mod):
    g = golden("lcg_uniform.npz")
    for seed in [0, 1, 2, 3, 290374]:
        lo, hi = g[f"s{seed}_dom"]
        ours = oracle_mod.lcg_uniform(seed, 1024, float(lo), float(hi))
        assert np.array_equal(ours, g[f"s{seed}"]), seed


def test_synth_generator_matches_reference(oracle_mod):
    from cubql_amd import synth
    g = golden("lcg_uniform.npz")
    for seed in [0, 1, 2, 3, 290374]:
        lo, hi = g[f"s{seed}_dom"]
        assert np.array_equal(synth.lcg_uniform(seed, 1024, float(lo), float(hi)), g[f"s{seed}"])


@pytest.mark.parametrize("case", CASES)
def test_sm_build_equals_reference_tree(oracle_mod, case):
    g = golden(f"sm_tree_{case}.npz")
    tree = oracle_mod.sm_build(g["boxes"], int(g["leaf_threshold"]))
    ref_nodes = g["nodes"].view(oracle_mod.NODE_DTYPE)
    assert len(tree.nodes) == len(ref_nodes)
    assert np.array_equal(tree.nodes["oc"], ref_nodes["oc"])
    assert np.array_equal(tree.nodes["lo"], ref_nodes["lo"])
    assert np.array_equal(tree.nodes["hi"], ref_nodes["hi"])
    assert np.array_equal(tree.prim_ids, g["prim_ids"])
    digest = hashlib.sha256(oracle_mod.canonical(tree).tobytes()).hexdigest()
    assert digest == str(g["digest"])
    code, _, _ = oracle_mod.check_tree(tree, g["boxes"], int(g["leaf_threshold"]))
    assert code == 0


@pytest.mark.parametrize("case", [c for c in CASES if c != "boxes_invalid"])
def test_fcp_knn_reference_mode_equals_reference(oracle_mod, case):
    g = golden(f"sm_tree_{case}.npz")
    tree = oracle_mod.Tree(g["nodes"].view(oracle_mod.NODE_DTYPE), g["prim_ids"])
    ids, d2 = oracle_mod.fcp(tree, g["points"], g["queries"], lowest_id=False)
    assert np.array_equal(ids, g["fcp_ids"])
    assert np.array_equal(d2, g["fcp_d2"])
    kid, kd, nf = oracle_mod.knn(tree, g["points"], g["queries"], 8, lowest_id=False)
    assert np.array_equal(kid, g["knn_ids"])
    assert np.array_equal(kd, g["knn_d2"])
    assert np.array_equal(nf, g["knn_found"])


@pytest.mark.parametrize("case", [c for c in CASES if c != "boxes_invalid"])
def test_lowest_id_mode_equals_brute_force(oracle_mod, case):
    g = golden(f"sm_tree_{case}.npz")
    tree = oracle_mod.Tree(g["nodes"].view(oracle_mod.NODE_DTYPE), g["prim_ids"])
    P, Q = g["points"], g["queries"][:1500]
    ids, d2 = oracle_mod.fcp(tree, P, Q, lowest_id=True)
    bi, bd = oracle_mod.fcp_brute(P, Q)
    assert np.array_equal(ids, bi)
    assert np.array_equal(d2, bd)
    kid, kd, _ = oracle_mod.knn(tree, P, Q, 8, lowest_id=True)
    bki, bkd = oracle_mod.knn_brute(P, Q, 8)
    assert np.array_equal(kid, bki)
    assert np.array_equal(kd, bkd)
    # against the reference's first-found rule: only exact ties may differ
    diff = ids != g["fcp_ids"][:1500]
    assert np.array_equal(d2, g["fcp_d2"][:1500])
    assert np.all(ids[diff] < g["fcp_ids"][:1500][diff])


def test_float_distance_vs_fp64(oracle_mod):
    g = golden("sm_tree_uniform8k.npz")
    P, Q = g["points"], g["queries"][:2000]
    fi, fd = oracle_mod.fcp_brute(P, Q)
    di, dd = oracle_mod.fcp_brute_f64(P, Q)
    # the float formula rounds 5 times: relative error <= ~3 ulp
    rel = np.abs(fd.astype(np.float64) - dd) / np.maximum(dd, 1e-300)
    assert rel.max() < 4 * 2.0 ** -24
    # ids agree except where the float rounding makes two candidates tie or swap
    mism = fi != di
    if mism.any():
        exact_f = ((P[fi[mism]].astype(np.float64) - Q[mism]) ** 2).sum(1)
        assert np.all(np.abs(exact_f - dd[mism]) <= 4 * 2.0 ** -24 * dd[mism])


def test_c1_tree_and_fcp_match_reference(oracle_mod):
    from cubql_amd import synth
    g = golden("c1_fcp.npz")
    P = synth.lcg_uniform(0, 100000, 0.0, 1.0)
    Q = synth.lcg_uniform(1, 100000, -0.1, 1.1)
    assert hashlib.sha256(P.tobytes()).hexdigest() == str(g["points_sha"])
    assert hashlib.sha256(Q.tobytes()).hexdigest() == str(g["queries_sha"])
    tree = oracle_mod.sm_build(oracle_mod.points_to_boxes(P), 8)
    assert len(tree.nodes) == int(g["num_nodes"])
    assert hashlib.sha256(oracle_mod.canonical(tree).tobytes()).hexdigest() == str(g["digest"])
    ids, d2 = oracle_mod.fcp(tree, P, Q, lowest_id=False)
    assert np.array_equal(ids, g["fcp_ids"])
    assert np.array_equal(d2, g["fcp_d2"])


def test_check_tree_catches_planted_mistakes(oracle_mod):
    g = golden("sm_tree_uniform8k.npz")
    boxes = g["boxes"]
    tree = oracle_mod.Tree(g["nodes"].view(oracle_mod.NODE_DTYPE).copy(), g["prim_ids"].copy())
    assert oracle_mod.check_tree(tree, boxes, 8)[0] == 0
    # 1. a shrunk inner box
    t1 = oracle_mod.Tree(tree.nodes.copy(), tree.prim_ids.copy())
    inner = np.nonzero(t1.counts == 0)[0][3]
    t1.nodes["lo"][inner][0] += np.float32(1e-3)
    assert oracle_mod.check_tree(t1, boxes, 8)[0] != 0
    # 2. a duplicated prim
    t2 = oracle_mod.Tree(tree.nodes.copy(), tree.prim_ids.copy())
    t2.prim_ids[5] = t2.prim_ids[6]
    assert oracle_mod.check_tree(t2, boxes, 8)[0] != 0
    # 3. leaf too large for the cap
    assert oracle_mod.check_tree(tree, boxes, 2)[0] != 0


def test_canonical_is_order_free_but_structure_sensitive(oracle_mod):
    g = golden("sm_tree_uniform8k.npz")
    tree = oracle_mod.Tree(g["nodes"].view(oracle_mod.NODE_DTYPE).copy(), g["prim_ids"].copy())
    base = oracle_mod.canonical(tree)
    # reverse each leaf's prim order and swap the root's children: same canonical form
    t = oracle_mod.Tree(tree.nodes.copy(), tree.prim_ids.copy())
    for i in np.nonzero(t.counts > 0)[0]:
        o, c = t.offsets[i], t.counts[i]
        t.prim_ids[o:o + c] = t.prim_ids[o:o + c][::-1].copy()
    r = int(t.offsets[0])
    a, b = t.nodes[r].copy(), t.nodes[r + 1].copy()
    t.nodes[r], t.nodes[r + 1] = b, a
    assert np.array_equal(oracle_mod.canonical(t), base)
    # swapping two prims between different leaves changes it
    leaves = np.nonzero(tree.counts > 0)[0]
    t3 = oracle_mod.Tree(tree.nodes.copy(), tree.prim_ids.copy())
    i, j = tree.offsets[leaves[0]], tree.offsets[leaves[-1]]
    t3.prim_ids[i], t3.prim_ids[j] = t3.prim_ids[j], t3.prim_ids[i]
    assert not np.array_equal(oracle_mod.canonical(t3), base)


def _spread_numpy(v, bits):
    r = np.zeros_like(v, dtype=np.uint64)
    for i in range(bits):
        r |= ((v.astype(np.uint64) >> np.uint64(i)) & np.uint64(1)) << np.uint64(3 * i)
    return r


@pytest.mark.parametrize("bits", [10, 21])
def test_morton_closed_form(oracle_mod, bits):
    # centroids placed exactly on cell corners of a [0,1)^3 grid scaled to 2^bits
    rng = np.random.default_rng(3)
    n = 4096
    cells = rng.integers(0, 1 << bits, size=(n, 3), dtype=np.int64)
    cells[0] = 0
    cells[1] = (1 << bits) - 1
    extent = np.float32(1 << bits)
    pts = cells.astype(np.float32)
    pts[0] = 0.0
    pts[1] = extent  # upper corner quantises to 2^b and must clamp to 2^b-1
    keys, ids = oracle_mod.morton_keys(oracle_mod.points_to_boxes(pts), bits)
    assert np.array_equal(ids, np.arange(n))
    c = cells.copy()
    c[1] = (1 << bits) - 1
    want = (_spread_numpy(c[:, 0], bits) | (_spread_numpy(c[:, 1], bits) << np.uint64(1))
            | (_spread_numpy(c[:, 2], bits) << np.uint64(2)))
    # scale = 2^b / extent == 1 exactly, so cells are the integer coordinates
    assert np.array_equal(keys, want)


def test_radix_hierarchy_rules(oracle_mod):
    from cubql_amd import synth
    P = synth.lcg_uniform(5, 20000)
    P[:300] = P[0]  # a block of equal keys
    boxes = oracle_mod.points_to_boxes(P)
    T = 4
    keys, ids = oracle_mod.morton_keys(boxes, 10)
    keys, ids = oracle_mod.sort_pairs(keys, ids)
    tree = oracle_mod.radix_build(boxes, 10, T)
    assert np.array_equal(tree.prim_ids, ids)
    assert oracle_mod.check_tree(tree, boxes, 0)[0] == 0
    # walk: each inner node's children split its range per radix.h:446-481
    stack = [(0, 0, len(keys))]
    while stack:
        nid, b, e = stack.pop()
        cnt, off = int(tree.counts[nid]), int(tree.offsets[nid])
        if cnt:
            assert (off, cnt) == (b, e - b) and cnt <= T
            continue
        assert e - b > T
        if keys[b] == keys[e - 1]:
            split = (b + e) // 2
        else:
            bit = (int(keys[b]) ^ int(keys[e - 1])).bit_length() - 1
            split = b + int(np.argmax((keys[b:e] >> np.uint64(bit)) & np.uint64(1)))
        stack.append((off, b, split))
        stack.append((off + 1, split, e))


def test_atomicbox_encoding_is_order_preserving(oracle_mod):
    L = oracle_mod.lib()
    vals = np.array([-np.inf, -3.4e38, -1.0, -1e-38, -1e-45, -0.0, 0.0, 1e-45, 1e-38, 1.0,
                     3.4e38, np.inf], dtype=np.float32)
    enc = [L.orc_atomicbox_encode(float(v)) for v in vals]
    assert all(a <= b for a, b in zip(enc, enc[1:]))
    assert enc[5] < enc[6]  # -0 sorts below +0
    for v, e in zip(vals, enc):
        assert np.float32(L.orc_atomicbox_decode(e)) == v
