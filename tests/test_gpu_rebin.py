"""GPU parity: the robust rebinning Morton builder (cuda::rebinRadixBuilder).

Through the C ABI, against the oracle's restatement (orc_rebin_build, pinned
in tests/test_oracle_rebin.py by an independent pure-Python version):

* primIDs identical position for position (Morton order after every broad
  rebin sort and every thread-phase exchange sort);
* identical canonical form (every split and leaf);
* TreeChecker: exact boxes, every valid prim once.

Cases: uniform points (no rebinning), tight clusters beside far outliers
(broad-phase rebins), exact duplicates (position keys), ranges around the
32-prim thread-phase limit, invalid boxes, and the C4 workload: 2,000,000
points of a 64-component Gaussian mixture with 1% outliers at 1e6, kNN k=16
and k=8 on a sample against the oracle, and the mean-kNN-distance backward against
the FP64 closed form (rule in tests/gradtol.py).
"""
import numpy as np
import pytest
import torch
from gradtol import check_point_grads

pytestmark = pytest.mark.gpu
F32 = np.float32
EPS = 2.0 ** -24


@pytest.fixture(scope="module")
def cq():
    import cubql_amd
    cubql_amd.lib()
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return cubql_amd


def _check_equal(cq, oracle_mod, boxes, T):
    bvh = cq.cuda.rebinRadixBuilder(torch.from_numpy(boxes).cuda(), cq.BuildConfig(makeLeafThreshold=T))
    torch.cuda.synchronize()
    nv = int((boxes[:, 0] <= boxes[:, 3]).sum())
    assert bvh.numPrims == nv
    tree = oracle_mod.Tree(cq.nodes_to_numpy(bvh).view(oracle_mod.NODE_DTYPE),
                           bvh.prim_ids_tensor().cpu().numpy().view(np.uint32))
    assert oracle_mod.check_tree(tree, boxes, max(T, 32))[0] == 0
    ref = oracle_mod.rebin_build(boxes, T)
    assert np.array_equal(tree.prim_ids, ref.prim_ids), "primIDs differ from the oracle"
    assert np.array_equal(oracle_mod.canonical(tree), oracle_mod.canonical(ref))
    return bvh, tree


def _clustered(n_per, centres, sigma, seed, outliers=()):
    rng = np.random.default_rng(seed)
    pts = [np.asarray(c, dtype=np.float64) + sigma * rng.standard_normal((n_per, 3)) for c in centres]
    pts.append(np.asarray(outliers, dtype=np.float64).reshape(-1, 3))
    p = np.concatenate(pts).astype(F32)
    return p[rng.permutation(len(p))]


@pytest.mark.parametrize("T", [1, 8])
def test_gpu_rebin_uniform(cq, oracle_mod, T):
    print(f"[gpu] rebin uniform T={T}", flush=True)
    bvh, _ = _check_equal(cq, oracle_mod, oracle_mod.points_to_boxes(oracle_mod.lcg_uniform(3, 100_000)), T)
    bvh.free()


@pytest.mark.parametrize("T,sigma", [(1, 1e-3), (4, 1e-4), (8, 1e-2)])
def test_gpu_rebin_clusters_and_outliers(cq, oracle_mod, T, sigma):
    print(f"[gpu] rebin clusters T={T} sigma={sigma}", flush=True)
    rng = np.random.default_rng(5)
    centres = rng.uniform(-1e3, 1e3, (40, 3))
    P = _clustered(2000, centres, sigma, seed=T, outliers=rng.uniform(-1e6, 1e6, (500, 3)))
    P = np.concatenate([P, np.tile(P[:1], (300, 1)), np.tile(P[7:8], (20, 1))])
    bvh, _ = _check_equal(cq, oracle_mod, oracle_mod.points_to_boxes(P), T)
    bvh.free()


@pytest.mark.parametrize("n", [1, 2, 31, 32, 33, 100, 5000])
def test_gpu_rebin_identical_points(cq, oracle_mod, n):
    print(f"[gpu] rebin identical n={n}", flush=True)
    B = np.tile(np.array([[1, 2, 3, 1, 2, 3]], dtype=F32), (n, 1))
    bvh, _ = _check_equal(cq, oracle_mod, B, 1)
    bvh.free()


def test_gpu_rebin_invalid_boxes(cq, oracle_mod):
    print("[gpu] rebin invalid boxes", flush=True)
    P = _clustered(3000, [[0, 0, 0], [10, 10, 10]], 1e-3, seed=9, outliers=[[1e6, 0, 0]])
    B = oracle_mod.points_to_boxes(P)
    B[::11, 0], B[::11, 3] = F32(1.0), F32(0.0)
    bvh, _ = _check_equal(cq, oracle_mod, B, 2)
    bvh.free()


def test_gpu_rebin_c4_workload(cq, oracle_mod):
    """2M-point Gaussian mixture with outliers, leaves <= 8; kNN k=16 and
    the backward of the mean kNN distance on a sample."""
    print("[gpu] rebin C4 full size", flush=True)
    from cubql_amd import synth
    P = synth.gaussian_mixture(2_000_000, seed=41)
    Q = synth.gaussian_mixture(2_000_000, seed=42)
    boxes = oracle_mod.points_to_boxes(P)
    bvh, tree = _check_equal(cq, oracle_mod, boxes, 8)
    dP, dQ = torch.from_numpy(P).cuda(), torch.from_numpy(Q).cuda()
    ids, d2, nf = cq.points.findKNN(bvh, dP, dQ, 16)
    g = torch.full((len(Q), 16), 1.0 / (len(Q) * 16), device="cuda")
    gq = torch.empty_like(dQ)
    gp = torch.zeros_like(dP)
    cq.points_backward(dP, dQ, ids, g, gq, gp)
    torch.cuda.synchronize()
    sel = np.arange(0, len(Q), len(Q) // 4000)
    oi, od, _ = oracle_mod.knn(tree, P, Q[sel], 16)
    ids_h = ids.cpu().numpy()
    assert np.array_equal(ids_h[sel], oi)
    assert np.array_equal(d2.cpu().numpy()[sel], od)
    # the adjoint of every query and every data point against the FP64
    # closed form at the kernel's ids (tolerance rule in tests/gradtol.py)
    check_point_grads(oracle_mod, P, Q, ids_h, np.full((len(Q), 16), 1.0 / (len(Q) * 16)),
                      gq.cpu().numpy(), gp.cpu().numpy(), label="C4 kNN16 mean")
    # k = 8 takes the seeded traversal on the LDS stack; the outlier queries
    # overflow it and rerun in the deep kernel
    ids8, d28, nf8 = cq.points.findKNN(bvh, dP, dQ, 8)
    torch.cuda.synchronize()
    oi8, od8, on8 = oracle_mod.knn(tree, P, Q[sel], 8)
    assert np.array_equal(ids8.cpu().numpy()[sel], oi8)
    assert np.array_equal(d28.cpu().numpy()[sel], od8)
    assert np.array_equal(nf8.cpu().numpy()[sel], on8)
    bvh.free()


@pytest.mark.parametrize("n,T", [(2000, 1), (2000, 8), (33, 1), (32, 1), (20, 2)])
def test_gpu_rebin_cluster_reduces_to_radix30(cq, oracle_mod, n, T):
    """An isolated cluster (one global Morton cell, distinct keys in its own
    centroid domain) is re-keyed in that domain (broad phase above 32 prims,
    thread phase at or below): its subtree in the GPU rebin tree equals the
    GPU radix30 tree built over the cluster alone, prim order included."""
    print(f"[gpu] rebin cluster reduction n={n} T={T}", flush=True)
    from test_oracle_rebin import (check_cluster_reduction, find_subtree, isolated_cluster,
                                   subtree_prim_order, subtree_tuple)
    P, cl = isolated_cluster(n, 100 + n + T)

    def gpu_tree(boxes, t, fn):
        bvh = fn(torch.from_numpy(boxes).cuda(), cq.BuildConfig(makeLeafThreshold=t))
        torch.cuda.synchronize()
        tree = oracle_mod.Tree(cq.nodes_to_numpy(bvh).view(oracle_mod.NODE_DTYPE),
                               bvh.prim_ids_tensor().cpu().numpy().view(np.uint32))
        bvh.free()
        return tree

    tree = check_cluster_reduction(oracle_mod, lambda b, t: gpu_tree(b, t, cq.cuda.rebinRadixBuilder),
                                   P, cl, T)
    alone = gpu_tree(oracle_mod.points_to_boxes(P[cl]), T, cq.cuda.radix30Builder)
    node = find_subtree(tree, cl)
    remap = {int(g): j for j, g in enumerate(cl)}
    ident = {j: j for j in range(len(cl))}
    assert subtree_prim_order(tree, node, remap) == alone.prim_ids.tolist()
    assert subtree_tuple(tree, node, remap) == subtree_tuple(alone, 0, ident)
