"""GPU parity of the wide BVH (WideBVH<float,3,W>, W = 4 and 8).

* cubql_collapse_wide_f3 on spatial-median, Morton and SAH trees equals the
  oracle's collapse (orc_wide_collapse, pinned in tests/test_oracle_wide.py
  by the closed form) of the same binary tree, record for record;
* the wide gpuBuilder equals that collapse of the binary gpuBuilder tree
  and keeps its primIDs;
* wide closest-point and kNN (k = 1, 8, 16) ids and squared distances equal
  the binary path's and the oracle's at the C1 size (100K points, 100K
  queries) and on 1M queries over the 10M C2 points.
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


def _tree(cq, oracle_mod, bvh):
    return oracle_mod.Tree(cq.nodes_to_numpy(bvh).view(oracle_mod.NODE_DTYPE),
                           bvh.prim_ids_tensor().cpu().numpy().view(np.uint32))


def _same_wide(cq, wide, ref):
    got = cq.wide_nodes_to_numpy(wide)
    assert got.shape == ref.shape
    assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))


@pytest.mark.parametrize("builder", ["sm", "radix30", "sah"])
@pytest.mark.parametrize("W", [4, 8])
def test_collapse_equals_oracle(cq, oracle_mod, builder, W):
    print(f"[gpu] wide collapse {builder} W={W}", flush=True)
    from cubql_amd import synth
    P = torch.from_numpy(synth.lcg_uniform(21, 150_000)).cuda()
    boxes = cq.points_to_boxes(P)
    cfg = cq.BuildConfig(makeLeafThreshold=4)
    bvh = {"sm": cq.gpuBuilder, "radix30": cq.cuda.radix30Builder, "sah": cq.cuda.sahBuilder}[builder](boxes, cfg)
    wide = cq.collapse_wide(bvh, W)
    torch.cuda.synchronize()
    tree = _tree(cq, oracle_mod, bvh)
    _same_wide(cq, wide, oracle_mod.wide_collapse(tree, W).view(oracle_mod.NODE_DTYPE))
    assert torch.equal(wide.prim_ids_tensor(), bvh.prim_ids_tensor())
    wide.free()
    bvh.free()


@pytest.mark.parametrize("W", [4, 8])
def test_wide_builder_and_queries_c1(cq, oracle_mod, W):
    print(f"[gpu] wide builder + queries C1 W={W}", flush=True)
    from cubql_amd import synth
    P = synth.lcg_uniform(0, 100_000)
    Q = synth.lcg_uniform(1, 100_000, -0.1, 1.1)
    dP, dQ = torch.from_numpy(P).cuda(), torch.from_numpy(Q).cuda()
    cfg = cq.BuildConfig(makeLeafThreshold=8)
    wide = cq.wideGpuBuilder(cq.points_to_boxes(dP), W, cfg)
    bvh = cq.gpuBuilder(cq.points_to_boxes(dP), cfg)
    torch.cuda.synchronize()
    _same_wide(cq, wide, oracle_mod.wide_collapse(_tree(cq, oracle_mod, bvh), W).view(oracle_mod.NODE_DTYPE))
    assert torch.equal(wide.prim_ids_tensor(), bvh.prim_ids_tensor())
    wi, wd = cq.points.findClosest(wide, dP, dQ)
    bi, bd = cq.points.findClosest(bvh, dP, dQ)
    torch.cuda.synchronize()
    assert torch.equal(wi, bi) and torch.equal(wd, bd)
    oi, od = oracle_mod.fcp(_tree(cq, oracle_mod, bvh), P, Q, lowest_id=True)
    assert np.array_equal(wi.cpu().numpy(), oi) and np.array_equal(wd.cpu().numpy(), od)
    for k in (1, 8, 16):
        a = cq.points.findKNN(wide, dP, dQ, k)
        b = cq.points.findKNN(bvh, dP, dQ, k)
        torch.cuda.synchronize()
        for x, y in zip(a, b):
            assert torch.equal(x, y), k
        ki, kd = oracle_mod.knn_brute(P, Q[:3000], k)
        assert np.array_equal(a[0].cpu().numpy()[:3000], ki) and np.array_equal(a[1].cpu().numpy()[:3000], kd)
    wide.free()
    bvh.free()


def test_wide_queries_c2(cq, oracle_mod):
    print("[gpu] wide queries C2 10M", flush=True)
    from cubql_amd import synth
    P = synth.lcg_uniform_torch(2, 10_000_000)
    Q = synth.lcg_uniform_torch(3, 1_000_000)
    cfg = cq.BuildConfig(makeLeafThreshold=8)
    bvh = cq.gpuBuilder(cq.points_to_boxes(P), cfg)
    for W in (4, 8):
        wide = cq.collapse_wide(bvh, W)
        wi, wd = cq.points.findClosest(wide, P, Q)
        bi, bd = cq.points.findClosest(bvh, P, Q)
        ka = cq.points.findKNN(wide, P, Q, 8)
        kb = cq.points.findKNN(bvh, P, Q, 8)
        torch.cuda.synchronize()
        assert torch.equal(wi, bi) and torch.equal(wd, bd), W
        for x, y in zip(ka, kb):
            assert torch.equal(x, y), W
        wide.free()
    bvh.free()
