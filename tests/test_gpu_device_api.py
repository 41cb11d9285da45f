"""GPU run of user kernels written against the reference-spelled device
headers (include/cuBQL/...): tests/cpp/device_api_sample.hip is shaped like
the reference's samples/s01_closestPoint_points_gpu/closestPoint.cu:31-66
(computeBoxes with box3f().including, gpuBuilder, a runQueries kernel that
calls cuBQL::points::findClosest per thread), its WideBVH form
(closestPoint_WideBVH.cu, width 8), plus points::findKNN,
shrinkingRadiusQuery::forEachLeaf, fixedBoxQuery::forEachPrim (binary and
wide) and triangles::findClosest / fcp kernels.

Its outputs are compared with the oracle: closest ids with the float brute
force (lowest id on ties), kNN rows with the brute-force kNN, in-radius
counts and in-box counts with brute-force counts, and triangle ids, squared distances and
closest points bit for bit with the oracle's triangle query on the tree
the program built.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_device_api_sample(oracle_mod, tmp_path):
    print("[gpu] device API sample program", flush=True)
    from cubql_amd import synth
    exe = os.path.join(ROOT, "cubql_amd", "device_api_sample")
    assert os.path.exists(exe), "run __graft_entry__.build() first"
    P = synth.lcg_uniform(290374, 10_000)
    Q = synth.lcg_uniform(1234567, 2_000)
    V, F = synth.noisy_icosphere(4, seed=3)
    Qt = synth.mesh_surface_queries(V, F, 2_000, seed=4)
    # the same queries drive the point and the triangle kernels
    Qall = np.concatenate([Q, Qt]).astype(np.float32)
    P.tofile(tmp_path / "data.f32")
    Qall.tofile(tmp_path / "queries.f32")
    V.tofile(tmp_path / "verts.f32")
    F.astype(np.int32).tofile(tmp_path / "tris.i32")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "OK device api" in r.stdout
    nq = len(Qall)
    rd = lambda name, dt: np.fromfile(tmp_path / name, dtype=dt)  # noqa: E731
    bi, _ = oracle_mod.fcp_brute(P, Qall)
    assert np.array_equal(rd("fcp.i32", np.int32), bi)
    ki, kd = oracle_mod.knn_brute(P, Qall, 8)
    assert np.array_equal(rd("knn_ids.i32", np.int32).reshape(nq, 8), ki)
    assert np.array_equal(rd("knn_d2.f32", np.float32).reshape(nq, 8), kd)
    assert np.all(rd("knn_found.i32", np.int32) == 8)
    cnt = rd("leaf_count.i32", np.int32)
    # the kernel counts with float32 (x*x+y*y)+z*z; compare in float32
    d2f = np.zeros((nq, len(P)), np.float32)
    for dim in range(3):
        diff = (P[None, :, dim] - Qall[:, None, dim]).astype(np.float32)
        d2f = (d2f + diff * diff).astype(np.float32)
    assert np.array_equal(cnt, (d2f <= np.float32(0.01)).sum(1))
    assert np.array_equal(rd("fcp_wide.i32", np.int32), bi)
    inbox = np.ones((nq, len(P)), bool)
    for dim in range(3):
        lo = (Qall[:, dim] - np.float32(0.05)).astype(np.float32)
        hi = (Qall[:, dim] + np.float32(0.05)).astype(np.float32)
        inbox &= (P[None, :, dim] >= lo[:, None]) & (P[None, :, dim] <= hi[:, None])
    assert np.array_equal(rd("box_count.i32", np.int32), inbox.sum(1))
    assert np.array_equal(rd("box_count_wide.i32", np.int32), inbox.sum(1))
    nodes = rd("tri_nodes.u8", np.uint8).view(oracle_mod.NODE_DTYPE)
    tree = oracle_mod.Tree(nodes, rd("tri_prims.u32", np.uint32))
    assert oracle_mod.check_tree(tree, oracle_mod.tri_boxes(F, V), 4)[0] == 0
    oi, od, oP = oracle_mod.fcp_tris(tree, F, V, Qall)
    assert np.array_equal(rd("tri_ids.i32", np.int32), oi)
    assert np.array_equal(rd("tri_d2.f32", np.float32).view(np.uint32), od.view(np.uint32))
    assert np.array_equal(rd("tri_P.f32", np.float32).reshape(nq, 3), oP)
    assert np.array_equal(rd("tri_ids_indexed.i32", np.int32), oi)
