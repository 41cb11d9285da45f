"""GPU parity: spatial-median gpuBuilder + closest-point / kNN queries.

Every call goes through the C ABI (libcubql_amd.so via cubql_amd).  Checks:

* the GPU tree equals the reference cpu::spatialMedian tree (golden
  fixtures written by the reference itself): identical primID array over the
  valid prims, identical canonical form (digest), exact boxes (TreeChecker);
* fcp ids/squared distances: bit-exact against the oracle's lowest-id mode,
  and against the reference's own results up to exact float ties (where the
  reference keeps the first found, we keep the lowest primID);
* kNN (k=8) likewise;
* at the 10M size: tree validity and query results on a sample, checked
  against the oracle (lowest-id rule) on the same tree.
"""
import hashlib
import math
import zlib

import numpy as np
import pytest
import torch

from conftest import golden

pytestmark = pytest.mark.gpu

CASES = ["uniform8k", "dupes", "identical", "adjacent", "boxes_invalid", "clustered"]


@pytest.fixture(scope="module")
def cq():
    import cubql_amd
    cubql_amd.lib()
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return cubql_amd


def _gpu_tree(cq, oracle_mod, bvh):
    nodes = cq.nodes_to_numpy(bvh)
    ids = bvh.prim_ids_tensor().cpu().numpy().view(np.uint32)
    return oracle_mod.Tree(nodes.view(oracle_mod.NODE_DTYPE), ids)


def _valid_count(boxes):
    return int((boxes[:, 0] <= boxes[:, 3]).sum())


@pytest.mark.parametrize("case", CASES)
def test_gpu_builder_equals_reference_tree(cq, oracle_mod, case):
    print(f"[gpu] sm tree case={case}", flush=True)
    g = golden(f"sm_tree_{case}.npz")
    boxes = g["boxes"]
    T = int(g["leaf_threshold"])
    bvh = cq.gpuBuilder(torch.from_numpy(boxes).cuda(), cq.BuildConfig(makeLeafThreshold=T))
    torch.cuda.synchronize()
    tree = _gpu_tree(cq, oracle_mod, bvh)
    nv = _valid_count(boxes)
    assert bvh.numPrims == boxes.shape[0]
    # exact boxes, every valid prim exactly once, leaves <= T
    code, depth, leaves = oracle_mod.check_tree(tree, boxes, T)
    assert code == 0, code
    # same canonical tree as the reference CPU builder
    assert hashlib.sha256(oracle_mod.canonical(tree).tobytes()).hexdigest() == str(g["digest"])
    # identical prim order over the valid prims (same partition order)
    assert np.array_equal(tree.prim_ids[:nv], g["prim_ids"])
    # invalid prims listed after all valid ones, ascending (sm_builder.h:590)
    inv = np.nonzero(~(boxes[:, 0] <= boxes[:, 3]))[0].astype(np.uint32)
    assert np.array_equal(tree.prim_ids[nv:], inv)
    # node 1 unused and zero (bvh.h:66-70)
    assert tree.nodes["oc"][1] == 0 if len(tree.nodes) > 1 else True
    bvh.free()


@pytest.mark.parametrize("case", [c for c in CASES if c != "boxes_invalid"])
def test_gpu_fcp_knn_parity(cq, oracle_mod, case):
    print(f"[gpu] fcp/knn case={case}", flush=True)
    g = golden(f"sm_tree_{case}.npz")
    P, Q = g["points"], g["queries"]
    T = int(g["leaf_threshold"])
    dP, dQ = torch.from_numpy(P).cuda(), torch.from_numpy(Q).cuda()
    bvh = cq.gpuBuilder(cq.points_to_boxes(dP), cq.BuildConfig(makeLeafThreshold=T))
    ids, d2 = cq.points.findClosest(bvh, dP, dQ)
    kid, kd, nf = cq.points.findKNN(bvh, dP, dQ, 8)
    torch.cuda.synchronize()
    ids, d2 = ids.cpu().numpy(), d2.cpu().numpy()
    kid, kd, nf = kid.cpu().numpy(), kd.cpu().numpy(), nf.cpu().numpy()
    tree = _gpu_tree(cq, oracle_mod, bvh)
    oi, od = oracle_mod.fcp(tree, P, Q, lowest_id=True)
    assert np.array_equal(ids, oi)
    assert np.array_equal(d2.view(np.uint32), od.view(np.uint32))
    # reference (first found): distances bit-exact, ids equal except exact ties
    assert np.array_equal(d2, g["fcp_d2"])
    diff = ids != g["fcp_ids"]
    assert np.all(ids[diff] < g["fcp_ids"][diff])
    oki, okd, onf = oracle_mod.knn(tree, P, Q, 8, lowest_id=True)
    assert np.array_equal(kid, oki)
    assert np.array_equal(kd.view(np.uint32), okd.view(np.uint32))
    assert np.array_equal(nf, onf)
    assert np.array_equal(kd, g["knn_d2"])
    bvh.free()


def test_gpu_c1_workload(cq, oracle_mod):
    """C1: 100K uniform points (seed 0), leaf <= 8, 100K queries (seed 1)."""
    print("[gpu] C1 100K/100K", flush=True)
    from cubql_amd import synth
    g = golden("c1_fcp.npz")
    P = synth.lcg_uniform(0, 100000, 0.0, 1.0)
    Q = synth.lcg_uniform(1, 100000, -0.1, 1.1)
    dP, dQ = torch.from_numpy(P).cuda(), torch.from_numpy(Q).cuda()
    bvh = cq.gpuBuilder(cq.points_to_boxes(dP), cq.BuildConfig(makeLeafThreshold=8))
    ids, d2 = cq.points.findClosest(bvh, dP, dQ)
    kid, kd, _ = cq.points.findKNN(bvh, dP, dQ[:4000], 8)
    torch.cuda.synchronize()
    tree = _gpu_tree(cq, oracle_mod, bvh)
    assert bvh.numNodes == int(g["num_nodes"]) + 1  # + the unused node 1
    assert hashlib.sha256(oracle_mod.canonical(tree).tobytes()).hexdigest() == str(g["digest"])
    ids, d2 = ids.cpu().numpy(), d2.cpu().numpy()
    assert np.array_equal(d2, g["fcp_d2"])
    diff = ids != g["fcp_ids"]
    assert np.all(ids[diff] < g["fcp_ids"][diff])
    oi, od = oracle_mod.fcp(tree, P, Q, lowest_id=True)
    assert np.array_equal(ids, oi)
    assert np.array_equal(kd.cpu().numpy(), g["knn_d2"])
    bvh.free()


def _deep_case(name):
    """Inputs that drive the two-level passes through their rarer paths: trees
    deeper than the size predicts (several host batches), one-sided splits of
    large nodes (position halves, reversed halves turned around), exact ties
    with the split planes, and boxes with invalid entries."""
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    T = 8
    if name == "skewed":            # density piles up near the origin: a deep tree
        P = rng.random((400_000, 3), dtype=np.float32) ** np.float32(6)
    elif name == "dupes_large":     # 40 distinct positions, ~7500 copies each
        table = rng.random((40, 3), dtype=np.float32)
        P = table[rng.integers(0, 40, 300_000)]
    elif name == "halfdup":         # half the points on one position
        P = rng.random((200_000, 3), dtype=np.float32)
        P[::2] = np.float32(0.25)
    elif name == "line":            # every split along x
        P = np.zeros((250_000, 3), np.float32)
        P[:, 0] = rng.random(250_000, dtype=np.float32)
    elif name == "signed":          # both signs on every axis, -0 and +0 included
        P = rng.random((300_000, 3), dtype=np.float32) * np.float32(2) - np.float32(1)
        P[::1000] = np.float32(0)
        P[1::1000] = np.float32(-0.0)
    elif name == "grid_ties":       # integer grid: planes fall on coordinates
        g = np.arange(64, dtype=np.float32)
        P = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
        P = P[rng.permutation(len(P))]
        T = 1
    else:                           # boxes, 1% of them invalid (lower > upper)
        lo = rng.random((200_000, 3), dtype=np.float32)
        ext = rng.random((200_000, 3), dtype=np.float32) * np.float32(0.01)
        boxes = np.concatenate([lo, lo + ext], 1)
        bad = rng.random(200_000) < 0.01
        boxes[bad, 0] = boxes[bad, 3] + np.float32(1)
        return boxes, T
    return np.concatenate([P, P], 1).astype(np.float32), T


@pytest.mark.parametrize("name", ["skewed", "dupes_large", "halfdup", "line", "signed", "grid_ties", "boxes"])
def test_gpu_builder_deep_and_degenerate_inputs(cq, oracle_mod, name):
    print(f"[gpu] sm deep case={name}", flush=True)
    boxes, T = _deep_case(name)
    bvh = cq.gpuBuilder(torch.from_numpy(boxes).cuda(), cq.BuildConfig(makeLeafThreshold=T))
    torch.cuda.synchronize()
    tree = _gpu_tree(cq, oracle_mod, bvh)
    nv = _valid_count(boxes)
    code, depth, leaves = oracle_mod.check_tree(tree, boxes, T)
    assert code == 0, code
    ref = oracle_mod.sm_build(boxes, T)
    # the CPU recursion's prim order and tree
    assert np.array_equal(tree.prim_ids[:nv], ref.prim_ids[:nv])
    assert np.array_equal(oracle_mod.canonical(tree), oracle_mod.canonical(ref))
    print(f"[gpu] sm deep case={name}: nodes={bvh.numNodes} depth={depth}", flush=True)
    bvh.free()


def test_gpu_empty_and_tiny_inputs(cq, oracle_mod):
    print("[gpu] tiny inputs", flush=True)
    # n == 0: untouched (gpu_builder.h:39)
    bvh = cq.gpuBuilder(torch.empty((0, 6), device="cuda"), cq.BuildConfig())
    assert bvh.numNodes == 0 and bvh.numPrims == 0
    for n in [1, 2, 3, 9, 257, 2049]:
        P = np.random.default_rng(n).random((n, 3), dtype=np.float32)
        dP = torch.from_numpy(P).cuda()
        bvh = cq.gpuBuilder(cq.points_to_boxes(dP), cq.BuildConfig(makeLeafThreshold=1))
        torch.cuda.synchronize()
        tree = _gpu_tree(cq, oracle_mod, bvh)
        assert oracle_mod.check_tree(tree, oracle_mod.points_to_boxes(P), 1)[0] == 0
        ref = oracle_mod.sm_build(oracle_mod.points_to_boxes(P), 1)
        assert np.array_equal(oracle_mod.canonical(tree), oracle_mod.canonical(ref))
        ids, d2 = cq.points.findClosest(bvh, dP, dP)
        assert np.array_equal(ids.cpu().numpy(), np.arange(n, dtype=np.int32))
        assert np.all(d2.cpu().numpy() == 0)
        # kNN with k up to 8 (the seeded traversal on the LDS stack), also
        # with fewer prims than k, against brute force; queries off the points
        Q = np.random.default_rng(n + 100).random((300, 3), dtype=np.float32)
        for k in (1, 4, 8):
            kid, kd, nf = cq.points.findKNN(bvh, dP, torch.from_numpy(Q).cuda(), k)
            bki, bkd = oracle_mod.knn_brute(P, Q, k)
            assert np.array_equal(kid.cpu().numpy(), bki), (n, k)
            assert np.array_equal(kd.cpu().numpy(), bkd), (n, k)
            assert np.all(nf.cpu().numpy() == min(n, k)), (n, k)
        bvh.free()
    # all-invalid input: no valid prims -> every query reports not found
    boxes = torch.tensor([[1, 1, 1, 0, 0, 0]] * 5, dtype=torch.float32, device="cuda")
    bvh = cq.gpuBuilder(boxes, cq.BuildConfig())
    assert bvh.numNodes == 0 and bvh.numPrims == 5
    ids, d2 = cq.points.findClosest(bvh, torch.zeros((5, 3), device="cuda"),
                                    torch.zeros((3, 3), device="cuda"))
    assert (ids.cpu() == -1).all() and torch.isinf(d2.cpu()).all()
    bvh.free()


def test_gpu_max_search_radius(cq, oracle_mod):
    print("[gpu] sqrMax", flush=True)
    g = golden("sm_tree_uniform8k.npz")
    P, Q = g["points"], g["queries"]
    dP, dQ = torch.from_numpy(P).cuda(), torch.from_numpy(Q).cuda()
    bvh = cq.gpuBuilder(cq.points_to_boxes(dP), cq.BuildConfig(makeLeafThreshold=8))
    tree = _gpu_tree(cq, oracle_mod, bvh)
    for r2 in [0.0, 1e-4, 2e-3]:
        ids, d2 = cq.points.findClosest(bvh, dP, dQ, r2)
        oi, od = oracle_mod.fcp(tree, P, Q, sqr_max=r2, lowest_id=True)
        assert np.array_equal(ids.cpu().numpy(), oi)
        kid, kd, nf = cq.points.findKNN(bvh, dP, dQ, 5, r2)
        oki, okd, onf = oracle_mod.knn(tree, P, Q, 5, sqr_max=r2, lowest_id=True)
        assert np.array_equal(kid.cpu().numpy(), oki)
        assert np.array_equal(nf.cpu().numpy(), onf)
    bvh.free()


@pytest.mark.parametrize("k", [1, 3, 16, 33, 64])
def test_gpu_knn_k_values(cq, oracle_mod, k):
    print(f"[gpu] knn k={k}", flush=True)
    g = golden("sm_tree_dupes.npz")  # many exact ties
    P, Q = g["points"], g["queries"]
    dP, dQ = torch.from_numpy(P).cuda(), torch.from_numpy(Q).cuda()
    bvh = cq.gpuBuilder(cq.points_to_boxes(dP), cq.BuildConfig(makeLeafThreshold=8))
    kid, kd, nf = cq.points.findKNN(bvh, dP, dQ, k)
    bki, bkd = oracle_mod.knn_brute(P, Q, k)
    assert np.array_equal(kid.cpu().numpy(), bki)
    assert np.array_equal(kd.cpu().numpy(), bkd)
    bvh.free()


def test_gpu_10m_build_and_query(cq, oracle_mod):
    """C2 size: 10M points (seed 2); validity + sampled query parity."""
    print("[gpu] 10M build", flush=True)
    from cubql_amd import synth
    n = 10_000_000
    P = synth.lcg_uniform(2, n)
    Q = synth.lcg_uniform(3, 20000)
    dP, dQ = torch.from_numpy(P).cuda(), torch.from_numpy(Q).cuda()
    bvh = cq.gpuBuilder(cq.points_to_boxes(dP), cq.BuildConfig(makeLeafThreshold=8))
    ids, d2 = cq.points.findClosest(bvh, dP, dQ)
    kid, kd, _ = cq.points.findKNN(bvh, dP, dQ[:5000], 8)
    torch.cuda.synchronize()
    tree = _gpu_tree(cq, oracle_mod, bvh)
    code, depth, leaves = oracle_mod.check_tree(tree, oracle_mod.points_to_boxes(P), 8)
    assert code == 0
    print(f"[gpu] 10M: nodes={bvh.numNodes} depth={depth} leaves={leaves}", flush=True)
    oi, od = oracle_mod.fcp(tree, P, Q, lowest_id=True)
    assert np.array_equal(ids.cpu().numpy(), oi)
    assert np.array_equal(d2.cpu().numpy(), od)
    oki, okd, _ = oracle_mod.knn(tree, P, Q[:5000], 8, lowest_id=True)
    assert np.array_equal(kid.cpu().numpy(), oki)
    # same tree as the CPU builder (canonical form)
    ref = oracle_mod.sm_build(oracle_mod.points_to_boxes(P), 8)
    assert np.array_equal(oracle_mod.canonical(tree), oracle_mod.canonical(ref))
    bvh.free()
