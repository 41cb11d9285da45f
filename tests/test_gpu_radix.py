"""GPU parity: onesweep radix sort and the Morton (radix) builders.

* radix sort: equals a stable argsort (numpy, kind="stable") for 32/64-bit
  keys, partial key widths, many duplicates, n not a multiple of the tile;
* radixBuilder (63-bit) / radix30Builder (30-bit): primIDs (Morton order,
  ties by primID) equal the oracle's restatement of radix.h exactly, the
  canonical tree equals the oracle's, every box is exact (TreeChecker), and
  fcp/kNN on the tree agree with brute force.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cq():
    import cubql_amd
    cubql_amd.lib()
    return cubql_amd


@pytest.mark.parametrize("bits,n,dup", [(32, 100_000, False), (30, 1_000_003, True),
                                        (64, 250_001, False), (63, 1_000_000, True),
                                        (17, 4096, True), (8, 5, False), (64, 1, False)])
def test_radix_sort_matches_stable_argsort(cq, bits, n, dup):
    print(f"[gpu] sort bits={bits} n={n}", flush=True)
    rng = np.random.default_rng(bits * 7 + n)
    hi = (1 << bits) if bits < 64 else None
    if bits <= 32:
        k = rng.integers(0, hi if not dup else 97, size=n, dtype=np.uint64).astype(np.uint32)
        kt = torch.from_numpy(k.view(np.int32)).cuda()
    else:
        k = rng.integers(0, np.iinfo(np.uint64).max if hi is None else hi, size=n, dtype=np.uint64)
        if dup:
            k = k % np.uint64(1009)
        kt = torch.from_numpy(k.view(np.int64)).cuda()
    v = np.arange(n, dtype=np.int32)
    vt = torch.from_numpy(v).cuda()
    cq.radix_sort_pairs(kt, vt, bits)
    torch.cuda.synchronize()
    order = np.argsort(k, kind="stable")
    got_k = kt.cpu().numpy().view(k.dtype)
    assert np.array_equal(got_k, k[order])
    assert np.array_equal(vt.cpu().numpy(), v[order])


def _tree(cq, oracle_mod, bvh):
    return oracle_mod.Tree(cq.nodes_to_numpy(bvh).view(oracle_mod.NODE_DTYPE),
                           bvh.prim_ids_tensor().cpu().numpy().view(np.uint32))


def _inputs(case):
    from cubql_amd import synth
    rng = np.random.default_rng(11)
    if case == "uniform":
        P = synth.lcg_uniform(0, 100_000)
        return np.concatenate([P, P], 1), P
    if case == "dupes":
        base = rng.random((50, 3), dtype=np.float32)
        P = base[rng.integers(0, 50, 20_000)]
        return np.concatenate([P, P], 1), P
    if case == "identical":
        # one key for every prim: the whole tree is one middle-split subtree
        P = np.tile(rng.random((1, 3), dtype=np.float32), (300_000, 1))
        return np.concatenate([P, P], 1), None
    if case == "manydupes":
        # 20K distinct positions, 10 prims each: thousands of equal-key
        # subtrees hanging under split nodes
        base = rng.random((20_000, 3), dtype=np.float32)
        P = base[np.repeat(np.arange(20_000), 10)]
        P = P[rng.permutation(len(P))]
        return np.concatenate([P, P], 1), None
    if case == "clustered":
        P = np.concatenate([rng.normal(0, 1e-4, (5000, 3)), rng.normal(5, 1, (5000, 3)),
                            rng.normal(-1e5, 1, (30, 3))]).astype(np.float32)
        return np.concatenate([P, P], 1), P
    if case == "boxes_invalid":
        lo = rng.random((30_000, 3), dtype=np.float32)
        bx = np.concatenate([lo, lo + rng.random((30_000, 3), dtype=np.float32) * np.float32(0.02)], 1)
        inv = rng.random(30_000) < 0.1
        bx[inv, 0], bx[inv, 3] = np.float32(2.0), np.float32(1.0)
        return bx.astype(np.float32), None
    raise KeyError(case)


@pytest.mark.parametrize("bits", [10, 21])
@pytest.mark.parametrize("case", ["uniform", "dupes", "manydupes", "identical", "clustered", "boxes_invalid"])
@pytest.mark.parametrize("T", [1, 8])
def test_radix_builder_equals_oracle(cq, oracle_mod, bits, case, T):
    print(f"[gpu] radix bits={bits} case={case} T={T}", flush=True)
    boxes, P = _inputs(case)
    build = cq.cuda.radixBuilder if bits == 21 else cq.cuda.radix30Builder
    bvh = build(torch.from_numpy(boxes).cuda(), cq.BuildConfig(makeLeafThreshold=T))
    torch.cuda.synchronize()
    tree = _tree(cq, oracle_mod, bvh)
    ref = oracle_mod.radix_build(boxes, bits, T)
    nv = int((boxes[:, 0] <= boxes[:, 3]).sum())
    assert bvh.numPrims == nv
    assert np.array_equal(tree.prim_ids, ref.prim_ids)
    assert oracle_mod.check_tree(tree, boxes, T)[0] == 0
    assert np.array_equal(oracle_mod.canonical(tree), oracle_mod.canonical(ref))
    assert bvh.numNodes == len(ref.nodes)
    if P is not None:
        Q = P[::97] + np.float32(1e-3)
        dP, dQ = torch.from_numpy(P).cuda(), torch.from_numpy(Q).cuda()
        ids, d2 = cq.points.findClosest(bvh, dP, dQ)
        bi, bd = oracle_mod.fcp_brute(P, Q)
        assert np.array_equal(ids.cpu().numpy(), bi)
        assert np.array_equal(d2.cpu().numpy(), bd)
        kid, kd, _ = cq.points.findKNN(bvh, dP, dQ, 8)
        bki, bkd = oracle_mod.knn_brute(P, Q, 8)
        assert np.array_equal(kid.cpu().numpy(), bki)
        assert np.array_equal(kd.cpu().numpy(), bkd)
    bvh.free()


def test_radix_builder_10m(cq, oracle_mod):
    """C2 size, 30-bit keys: primIDs and canonical tree equal the oracle."""
    print("[gpu] radix30 10M", flush=True)
    from cubql_amd import synth
    P = synth.lcg_uniform(2, 10_000_000)
    boxes = np.concatenate([P, P], 1)
    dP = torch.from_numpy(P).cuda()
    bvh = cq.cuda.radix30Builder(cq.points_to_boxes(dP), cq.BuildConfig(makeLeafThreshold=8))
    torch.cuda.synchronize()
    tree = _tree(cq, oracle_mod, bvh)
    ref = oracle_mod.radix_build(boxes, 10, 8)
    assert np.array_equal(tree.prim_ids, ref.prim_ids)
    assert np.array_equal(oracle_mod.canonical(tree), oracle_mod.canonical(ref))
    assert oracle_mod.check_tree(tree, boxes, 8)[0] == 0
    Q = synth.lcg_uniform(3, 20_000)
    ids, d2 = cq.points.findClosest(bvh, dP, torch.from_numpy(Q).cuda())
    oi, od = oracle_mod.fcp(tree, P, Q, lowest_id=True)
    assert np.array_equal(ids.cpu().numpy(), oi)
    bvh.free()


def test_radix_tiny_and_empty(cq, oracle_mod):
    print("[gpu] radix tiny", flush=True)
    bvh = cq.cuda.radixBuilder(torch.empty((0, 6), device="cuda"), cq.BuildConfig())
    assert bvh.numNodes == 0
    for n in [1, 2, 3, 17]:
        P = np.random.default_rng(n).random((n, 3), dtype=np.float32)
        boxes = np.concatenate([P, P], 1)
        bvh = cq.cuda.radix30Builder(torch.from_numpy(boxes).cuda(), cq.BuildConfig(makeLeafThreshold=1))
        torch.cuda.synchronize()
        tree = _tree(cq, oracle_mod, bvh)
        ref = oracle_mod.radix_build(boxes, 10, 1)
        assert np.array_equal(oracle_mod.canonical(tree), oracle_mod.canonical(ref))
        assert oracle_mod.check_tree(tree, boxes, 1)[0] == 0
        bvh.free()
    boxes = torch.tensor([[1, 1, 1, 0, 0, 0]] * 4, dtype=torch.float32, device="cuda")
    bvh = cq.cuda.radixBuilder(boxes, cq.BuildConfig())
    assert bvh.numNodes == 0 and bvh.numPrims == 0
