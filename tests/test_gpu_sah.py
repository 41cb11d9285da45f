"""GPU parity: the binned SAH builder (cuda::sahBuilder / gpuBuilder SAH).

Every call goes through the C ABI.  The oracle (orc_sah_build, pinned in
tests/test_oracle_sah.py by brute-force optimality at every node) partitions
in the same order as the GPU (left stable, right reversed), so the checks are
exact:

* primIDs over the valid prims are identical position for position;
* the canonical forms (children ordered by min primID, leaf sets sorted) are
  identical, so every split and leaf decision matches;
* TreeChecker: exact boxes, every valid prim once, leaves <= the cap.

Cases: random boxes at several bin counts and leaf thresholds, sizes around
the wave-finished node limit (256), identical boxes (a leaf, or middle splits
under a leaf cap; above 256 prims this happens in the large-node phase),
invalid boxes, a triangle mesh, and the full 1,310,720-triangle icosphere at
32 bins with leaves <= 4.
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
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return cubql_amd


def _random_boxes(n, seed, half=0.02):
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3), dtype=np.float32)
    h = (rng.random((n, 3), dtype=np.float32) * F32(half)).astype(F32)
    return np.concatenate([c - h, c + h], axis=1).astype(F32)


def _check_equal(cq, oracle_mod, boxes, T, nb, cap, via_gpu_builder=False):
    cfg = cq.BuildConfig(makeLeafThreshold=T, maxAllowedLeafSize=cap)
    dB = torch.from_numpy(boxes).cuda()
    if via_gpu_builder:
        assert nb == 16
        bvh = cq.gpuBuilder(dB, cfg.enableSAH())
    else:
        bvh = cq.cuda.sahBuilder(dB, cfg, numBins=nb)
    torch.cuda.synchronize()
    nodes = cq.nodes_to_numpy(bvh).view(oracle_mod.NODE_DTYPE)
    ids = bvh.prim_ids_tensor().cpu().numpy().view(np.uint32)
    tree = oracle_mod.Tree(nodes, ids)
    nv = int((boxes[:, 0] <= boxes[:, 3]).sum())
    assert bvh.numPrims == len(boxes)
    code, depth, leaves = oracle_mod.check_tree(tree, boxes, min(cap, 65535))
    assert code == 0, code
    ref = oracle_mod.sah_build(boxes, T, nb, cap)
    assert np.array_equal(ids[:nv], ref.prim_ids), "primIDs differ from the oracle"
    assert sorted(ids[nv:].tolist()) == np.nonzero(boxes[:, 0] > boxes[:, 3])[0].tolist()
    assert np.array_equal(oracle_mod.canonical(tree), oracle_mod.canonical(ref))
    bvh.free()
    return tree, depth, leaves


@pytest.mark.parametrize("n,nb,T", [(20000, 16, 1), (20000, 32, 4), (50000, 8, 2), (30000, 64, 8)])
def test_gpu_sah_random_boxes(cq, oracle_mod, n, nb, T):
    print(f"[gpu] sah random n={n} nb={nb} T={T}", flush=True)
    _check_equal(cq, oracle_mod, _random_boxes(n, seed=n + nb + T), T, nb, 1 << 15)


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 255, 256, 257, 300, 2049, 8193])
def test_gpu_sah_sizes(cq, oracle_mod, n):
    print(f"[gpu] sah size n={n}", flush=True)
    _check_equal(cq, oracle_mod, _random_boxes(n, seed=n), 1, 16, 1 << 15)
    _check_equal(cq, oracle_mod, _random_boxes(n, seed=n), 4, 32, 4)


@pytest.mark.parametrize("count,cap", [(100, 1 << 15), (100, 4), (5000, 1 << 15), (5000, 4),
                                       (70000, 1 << 15)])
def test_gpu_sah_identical_boxes(cq, oracle_mod, count, cap):
    print(f"[gpu] sah identical count={count} cap={cap}", flush=True)
    B = np.tile(np.array([[0.25, 0.5, 0.75, 1.0, 1.5, 2.0]], dtype=F32), (count, 1))
    tree, depth, leaves = _check_equal(cq, oracle_mod, B, 1, 16, cap)
    if cap < count:
        assert tree.counts.max() <= cap
    else:
        assert leaves == 1 and tree.counts[0] == count
    if count > 65535:
        # the 16-bit count field: the cap stays at 65535
        assert tree.counts.max() <= 65535 and leaves >= 2


def test_gpu_sah_mixed_duplicates_and_invalid(cq, oracle_mod):
    print("[gpu] sah duplicates + invalid", flush=True)
    B = np.concatenate([_random_boxes(20000, seed=31), np.tile(_random_boxes(1, seed=32), (700, 1))])
    B[::13, 0], B[::13, 3] = F32(1.0), F32(0.0)  # lower.x > upper.x: invalid
    _check_equal(cq, oracle_mod, B, 1, 16, 1 << 15)
    _check_equal(cq, oracle_mod, B, 4, 32, 4)


def test_gpu_sah_points(cq, oracle_mod):
    print("[gpu] sah points + fcp", flush=True)
    P = oracle_mod.lcg_uniform(0, 100_000)
    Q = oracle_mod.lcg_uniform(1, 20_000, -0.1, 1.1)
    boxes = oracle_mod.points_to_boxes(P)
    _check_equal(cq, oracle_mod, boxes, 8, 16, 1 << 15, via_gpu_builder=True)
    bvh = cq.gpuBuilder(torch.from_numpy(boxes).cuda(), cq.BuildConfig(makeLeafThreshold=8).enableSAH())
    ids, d2 = cq.points.findClosest(bvh, torch.from_numpy(P).cuda(), torch.from_numpy(Q).cuda())
    bi, bd = oracle_mod.fcp_brute(P, Q)
    assert np.array_equal(ids.cpu().numpy(), bi) and np.array_equal(d2.cpu().numpy(), bd)
    bvh.free()


def test_gpu_sah_triangles_small_mesh(cq, oracle_mod):
    print("[gpu] sah triangle mesh", flush=True)
    from cubql_amd import synth
    V, F = synth.noisy_icosphere(5, seed=6)
    _check_equal(cq, oracle_mod, oracle_mod.tri_boxes(F, V), 4, 32, 4)


def test_gpu_sah_c3_full_size(cq, oracle_mod):
    """1,310,720 noisy-icosphere triangles, 32 bins, leaves <= 4; then
    closest-triangle queries on a 20K sample against the oracle."""
    print("[gpu] sah C3 full size", flush=True)
    from cubql_amd import synth
    V, F = synth.noisy_icosphere(8, seed=4)
    boxes = oracle_mod.tri_boxes(F, V)
    tree, depth, leaves = _check_equal(cq, oracle_mod, boxes, 4, 32, 4)
    print(f"[gpu] C3 SAH tree: depth {depth}, {leaves} leaves", flush=True)
    Q = synth.mesh_surface_queries(V, F, 20000, seed=5)
    dV, dF = torch.from_numpy(V).cuda(), torch.from_numpy(F).cuda()
    bvh = cq.cuda.sahBuilder(cq.triangles.boxes(dF, dV),
                             cq.BuildConfig(makeLeafThreshold=4, maxAllowedLeafSize=4), numBins=32)
    ids, d2, _ = cq.triangles.findClosest(bvh, dF, dV, torch.from_numpy(Q).cuda())
    oi, od, _ = oracle_mod.fcp_tris(tree, F, V, Q)
    assert np.array_equal(ids.cpu().numpy(), oi) and np.array_equal(d2.cpu().numpy(), od)
    bvh.free()


def test_gpu_sah_rejects_bad_bin_count(cq):
    with pytest.raises(cq.CubqlError):
        cq.cuda.sahBuilder(torch.zeros(10, 6, device="cuda"), None, numBins=65)
