"""Traversal statistics (cubql_query_stats_points_f3 / _triangles_f3), the
counters behind the query rooflines.

Closed forms:
* a tree that is one leaf of n points: closest-point and kNN queries load
  no node pair and test n prims each (the seeding descent visits the leaf;
  the traversal then leaves it out);
* a two-leaf tree (n in (T, 2T]): every query loads the root's child pair
  at least once.
Bounds on a real tree: at least one prim tested per query, node bytes a
multiple of 32, the query count returned.
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


def test_stats_single_leaf(cq):
    rng = np.random.default_rng(3)
    P = torch.from_numpy(rng.random((6, 3), dtype=np.float32)).cuda()
    Q = torch.from_numpy(rng.random((100, 3), dtype=np.float32)).cuda()
    bvh = cq.gpuBuilder(cq.points_to_boxes(P), cq.BuildConfig(makeLeafThreshold=8))
    st = cq.query_stats(bvh, Q, data=P, k=0)
    assert st == {"node_bytes": 0, "prims": 6 * 100, "queries": 100}
    st = cq.query_stats(bvh, Q, data=P, k=4)
    assert st == {"node_bytes": 0, "prims": 6 * 100, "queries": 100}
    bvh.free()


def test_stats_two_leaves_and_bounds(cq):
    from cubql_amd import synth
    P = torch.from_numpy(synth.lcg_uniform(0, 12)).cuda()
    Q = torch.from_numpy(synth.lcg_uniform(1, 50)).cuda()
    bvh = cq.gpuBuilder(cq.points_to_boxes(P), cq.BuildConfig(makeLeafThreshold=8))
    st = cq.query_stats(bvh, Q, data=P, k=0)
    assert st["queries"] == 50 and st["node_bytes"] >= 2 * 64 * 50 and st["node_bytes"] % 32 == 0
    bvh.free()
    P = torch.from_numpy(synth.lcg_uniform(2, 200_000)).cuda()
    Q = torch.from_numpy(synth.lcg_uniform(3, 20_000)).cuda()
    bvh = cq.gpuBuilder(cq.points_to_boxes(P), cq.BuildConfig(makeLeafThreshold=8))
    for k in (0, 1, 8, 16):
        st = cq.query_stats(bvh, Q, data=P, k=k)
        print(f"[gpu] stats k={k}: {st['node_bytes'] / 2e4:.0f} node B/query, "
              f"{st['prims'] / 2e4:.1f} prims/query", flush=True)
        assert st["queries"] == 20_000
        assert st["prims"] >= 20_000 * max(k, 1)
        assert st["node_bytes"] % 32 == 0 and st["node_bytes"] > 0
    bvh.free()


def test_stats_triangles(cq):
    from cubql_amd import synth
    V, F = synth.noisy_icosphere(3, seed=4)
    Q = synth.mesh_surface_queries(V, F, 5000, seed=5)
    dV, dF, dQ = (torch.from_numpy(V).cuda(), torch.from_numpy(F).cuda(), torch.from_numpy(Q).cuda())
    boxes = cq.triangles.boxes(dF, dV)
    bvh = cq.cuda.sahBuilder(boxes, cq.BuildConfig(makeLeafThreshold=4, maxAllowedLeafSize=4), numBins=32)
    st = cq.query_stats(bvh, dQ, tris=dF, verts=dV)
    assert st["queries"] == 5000 and st["prims"] >= 5000 and st["node_bytes"] % 32 == 0
    bvh.free()
