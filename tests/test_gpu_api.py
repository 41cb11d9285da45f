"""GPU checks of API-level behaviour around the hot path.

* torch.autograd.gradcheck in float64 of the test-only autograd Functions
  (tests/autograd_binding.py) that put the HIP adjoint kernels behind
  torch's autograd: point distances (closest point and kNN) and
  point-to-triangle distances, against float64 finite differences;
* cubql_refit_f3 (cuBQL::refit, sm_builder.h:616-693): node boxes zeroed
  and refitted must equal the builder's exact boxes, for spatial-median,
  Morton and SAH trees and for a tree numbered by the CPU builder (root
  children at nodes 1 and 2);
* a build on a stream other than torch's current one: its temporaries come
  from torch's caching allocator on that stream, so allocations and writes
  on the current stream right after the call cannot reuse them while the
  build still runs.
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


def test_gradcheck_point_distance(cq):
    print("[gpu] gradcheck point distance", flush=True)
    from autograd_binding import PointDistance
    g = torch.Generator().manual_seed(3)
    P = torch.rand((300, 3), generator=g, dtype=torch.float64).cuda()
    Q = torch.rand((40, 3), generator=g, dtype=torch.float64).cuda()
    # float32-representable inputs: the HIP ids refer to exactly these points
    P, Q = P.float().double().requires_grad_(), Q.float().double().requires_grad_()
    for k in (1, 8):
        fn = lambda q, p: PointDistance.apply(q, p, k)[0]  # noqa: E731
        assert torch.autograd.gradcheck(fn, (Q, P), eps=1e-6, atol=1e-6, rtol=1e-5,
                                        nondet_tol=1e-6, fast_mode=False)


def test_gradcheck_triangle_distance(cq):
    print("[gpu] gradcheck triangle distance", flush=True)
    from autograd_binding import TriangleDistance
    from cubql_amd import synth
    V, F = synth.noisy_icosphere(1, seed=8)
    Q = synth.mesh_surface_queries(V, F, 30, seed=9, sigma=0.1)
    dV = torch.from_numpy(V).double().cuda().requires_grad_()
    dQ = torch.from_numpy(Q).double().cuda().requires_grad_()
    dF = torch.from_numpy(F).cuda()
    d, _ = TriangleDistance.apply(dQ, dV, dF)
    assert float(d.min()) > 1e-3  # away from the surface: the distance is smooth
    fn = lambda q, v: TriangleDistance.apply(q, v, dF)[0]  # noqa: E731
    assert torch.autograd.gradcheck(fn, (dQ, dV), eps=1e-6, atol=1e-6, rtol=1e-5,
                                    nondet_tol=1e-6, fast_mode=False)


def _tree(cq, oracle_mod, bvh):
    return oracle_mod.Tree(cq.nodes_to_numpy(bvh).view(oracle_mod.NODE_DTYPE),
                           bvh.prim_ids_tensor().cpu().numpy().view(np.uint32))


@pytest.mark.parametrize("builder", ["sm", "radix30", "sah", "sm_boxes"])
def test_refit_restores_exact_boxes(cq, oracle_mod, builder):
    print(f"[gpu] refit {builder}", flush=True)
    from cubql_amd import synth
    if builder == "sm_boxes":
        V, F = synth.noisy_icosphere(5, seed=2)
        boxes = cq.triangles.boxes(torch.from_numpy(F).cuda(), torch.from_numpy(V).cuda())
    else:
        boxes = cq.points_to_boxes(torch.from_numpy(synth.lcg_uniform(5, 200_000)).cuda())
    cfg = cq.BuildConfig(makeLeafThreshold=4)
    bvh = {"sm": cq.gpuBuilder, "sm_boxes": cq.gpuBuilder, "radix30": cq.cuda.radix30Builder,
           "sah": cq.cuda.sahBuilder}[builder](boxes, cfg)
    torch.cuda.synchronize()
    before = cq.nodes_to_numpy(bvh).copy()
    nt = bvh.nodes_tensor()
    nt[:, :6] = 0.0
    torch.cuda.synchronize()
    cq.refit(bvh, boxes)
    torch.cuda.synchronize()
    after = cq.nodes_to_numpy(bvh)
    reach = _reachable(after)
    assert np.array_equal(after["lo"][reach], before["lo"][reach])
    assert np.array_equal(after["hi"][reach], before["hi"][reach])
    assert oracle_mod.check_tree(_tree(cq, oracle_mod, bvh), boxes.cpu().numpy())[0] == 0
    bvh.free()


def _reachable(nodes):
    oc = nodes["oc"].astype(np.uint64)
    off = (oc & np.uint64((1 << 48) - 1)).astype(np.int64)
    cnt = (oc >> np.uint64(48)).astype(np.int64)
    seen = np.zeros(len(nodes), bool)
    stack = [0]
    while stack:
        i = stack.pop()
        seen[i] = True
        if cnt[i] == 0:
            stack += [off[i], off[i] + 1]
    return seen


def test_refit_cpu_numbered_tree(cq, oracle_mod):
    """A tree from the CPU builder (root-first numbering, node 1 in use)
    uploaded as a BinaryBVH and refitted from zeroed boxes."""
    print("[gpu] refit cpu-numbered tree", flush=True)
    P = oracle_mod.lcg_uniform(6, 50_000)
    B = oracle_mod.points_to_boxes(P)
    ref = oracle_mod.sm_build(B, 2)
    assert int(ref.offsets[0]) == 1 and int(ref.counts[0]) == 0
    mem = cq.TorchMemory()
    nodes = torch.from_numpy(ref.nodes.view(np.uint8).copy()).cuda()
    ids = torch.from_numpy(ref.prim_ids.view(np.uint8).copy()).cuda()
    bvh = cq.BinaryBVH(mem)
    mem.live[nodes.data_ptr()] = nodes
    mem.live[ids.data_ptr()] = ids
    bvh.c.nodes, bvh.c.numNodes = nodes.data_ptr(), len(ref.nodes)
    bvh.c.primIDs, bvh.c.numPrims = ids.data_ptr(), len(ref.prim_ids)
    bvh.nodes_tensor()[:, :6] = 0.0
    cq.refit(bvh, torch.from_numpy(B).cuda())
    torch.cuda.synchronize()
    after = cq.nodes_to_numpy(bvh)
    assert np.array_equal(after["lo"], ref.nodes["lo"])
    assert np.array_equal(after["hi"], ref.nodes["hi"])


def test_build_on_non_current_stream(cq, oracle_mod):
    print("[gpu] build on a side stream", flush=True)
    P = oracle_mod.lcg_uniform(7, 2_000_000)
    boxes = cq.points_to_boxes(torch.from_numpy(P).cuda())
    ref = cq.gpuBuilder(boxes, cq.BuildConfig(makeLeafThreshold=8))
    torch.cuda.synchronize()
    ref_nodes, ref_ids = cq.nodes_to_numpy(ref), ref.prim_ids_tensor().cpu().numpy()
    side = torch.cuda.Stream()
    mem = cq.TorchMemory()
    for rep in range(3):
        side.wait_stream(torch.cuda.current_stream())
        bvh = cq.gpuBuilder(boxes, cq.BuildConfig(makeLeafThreshold=8), stream=side, mem=mem)
        # work on the current stream while the side stream may still run
        junk = [torch.full((1 << 22,), 7, dtype=torch.int32, device="cuda") for _ in range(8)]
        del junk
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        assert np.array_equal(cq.nodes_to_numpy(bvh).view(np.uint8), ref_nodes.view(np.uint8)), rep
        assert np.array_equal(bvh.prim_ids_tensor().cpu().numpy(), ref_ids), rep
        bvh.free(stream=side)
    ref.free()
