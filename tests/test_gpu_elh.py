"""GPU parity: the edge-length-heuristic builder (gpuBuilder with
buildMethod ELH, builder/cuda/elh_builder.h).

Through the C ABI (cubql_gpu_builder_f3).  The oracle (orc_elh_build, pinned
in tests/test_oracle_elh.py by brute-force optimality at every node)
partitions in the GPU's order, so the checks are exact: primIDs position for
position, canonical forms (every split and leaf decision), TreeChecker
(exact boxes, every valid prim once, leaves <= the cap), and closest-point
queries over the tree equal brute force.

Cases: random boxes and points at several leaf thresholds, sizes around the
wave-finished node limit (256 prims), identical boxes (one leaf, or middle
splits under a leaf cap), invalid boxes, tightly clustered centroids, and a
triangle mesh.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
F32 = np.float32


@pytest.fixture(scope="module")
def cq():
    import cubql_amd
    cubql_amd.lib()
    return cubql_amd


def _random_boxes(n, seed, half=0.02):
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3), dtype=np.float32)
    h = (rng.random((n, 3), dtype=np.float32) * F32(half)).astype(F32)
    return np.concatenate([c - h, c + h], axis=1).astype(F32)


def _check_equal(cq, oracle_mod, boxes, T, cap=1 << 15):
    cfg = cq.BuildConfig(makeLeafThreshold=T, maxAllowedLeafSize=cap).enableELH()
    bvh = cq.gpuBuilder(torch.from_numpy(boxes).cuda(), cfg)
    torch.cuda.synchronize()
    nodes = cq.nodes_to_numpy(bvh).view(oracle_mod.NODE_DTYPE)
    ids = bvh.prim_ids_tensor().cpu().numpy().view(np.uint32)
    tree = oracle_mod.Tree(nodes, ids)
    nv = int((boxes[:, 0] <= boxes[:, 3]).sum())
    code, depth, leaves = oracle_mod.check_tree(tree, boxes, min(cap, 65535))
    assert code == 0, code
    ref = oracle_mod.elh_build(boxes, T, cap)
    assert np.array_equal(ids[:nv], ref.prim_ids), "primIDs differ from the oracle"
    assert np.array_equal(oracle_mod.canonical(tree), oracle_mod.canonical(ref))
    bvh.free()
    return tree


@pytest.mark.parametrize("n,T,half", [(20000, 1, 0.02), (50000, 4, 0.0), (30000, 8, 0.1)])
def test_gpu_elh_random(cq, oracle_mod, n, T, half):
    print(f"[gpu] elh random n={n} T={T}", flush=True)
    _check_equal(cq, oracle_mod, _random_boxes(n, seed=n + T, half=half), T)


@pytest.mark.parametrize("n", [1, 2, 9, 255, 256, 257, 2049])
def test_gpu_elh_sizes(cq, oracle_mod, n):
    print(f"[gpu] elh size n={n}", flush=True)
    _check_equal(cq, oracle_mod, _random_boxes(n, seed=7 * n), 1)


def test_gpu_elh_degenerate(cq, oracle_mod):
    print("[gpu] elh degenerate", flush=True)
    same = np.tile(np.array([[0, 0, 0, 1, 1, 1]], dtype=F32), (5000, 1))
    t = _check_equal(cq, oracle_mod, same, 1)
    assert t.counts[0] == 5000  # the root is one leaf (node 1 stays unused)
    _check_equal(cq, oracle_mod, same, 1, cap=8)
    B = _random_boxes(3000, seed=3)
    B[::7, 0], B[::7, 3] = F32(1.0), F32(0.0)   # invalid
    _check_equal(cq, oracle_mod, B, 2)
    # centroids spanning ~1e-31: binned without an epsilon
    x = (F32(1.0) + np.arange(400, dtype=F32) * F32(2.0 ** -23)) * F32(1e-25)
    C = np.zeros((400, 6), F32)
    C[:, 0] = C[:, 3] = x
    t = _check_equal(cq, oracle_mod, C, 1)
    assert t.counts[0] == 0


def test_gpu_elh_points_and_mesh(cq, oracle_mod):
    print("[gpu] elh points and mesh", flush=True)
    from cubql_amd import synth
    P = synth.lcg_uniform(0, 100_000)
    Q = synth.lcg_uniform(1, 20_000, -0.1, 1.1)
    dP = torch.from_numpy(P).cuda()
    bvh = cq.gpuBuilder(cq.points_to_boxes(dP), cq.BuildConfig(makeLeafThreshold=8).enableELH())
    ids, d2 = cq.points.findClosest(bvh, dP, torch.from_numpy(Q).cuda())
    bi, bd = oracle_mod.fcp_brute(P, Q)
    assert np.array_equal(ids.cpu().numpy(), bi) and np.array_equal(d2.cpu().numpy(), bd)
    bvh.free()
    _check_equal(cq, oracle_mod, oracle_mod.points_to_boxes(P[:30000]), 8)
    V, F = synth.noisy_icosphere(4, seed=4)
    _check_equal(cq, oracle_mod, oracle_mod.tri_boxes(F, V), 4, cap=4)
