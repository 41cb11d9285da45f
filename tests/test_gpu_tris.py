"""GPU parity: closest point on triangle meshes and the distance adjoints.

* fcp over triangles: triangle ids, squared distances and closest points are
  bit-exact against the oracle's restatement of queries/triangles/
  findClosest.h on the same tree (indexed and non-indexed layouts);
  distances are compared with an FP64 brute force and the error is printed;
* backward (triangles and points): compared with the FP64 closed-form
  gradient (oracle; itself pinned by finite differences in
  tests/test_oracle_tris.py).  Triangles: |g - g64| <= 1e-5 |g64| +
  16 * 2^-24 * s / d per query.  The closest point P is computed from q
  (plane projection, edge clamp), so its coordinates carry float32 rounding
  of their own magnitude s, and q - P is off by ~2^-24 s absolute: relative
  to d that is s / d (written out in _grad_tol).  Points: q - p is formed
  from two inputs and is correctly rounded relative to itself, so the rule
  is 1e-5 of the gradient's own size with no s / d term (tests/gradtol.py).
"""
import numpy as np
import pytest
import torch
from gradtol import check_point_grads

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -24


@pytest.fixture(scope="module")
def cq():
    import cubql_amd
    cubql_amd.lib()
    return cubql_amd


def _tree(cq, oracle_mod, bvh):
    return oracle_mod.Tree(cq.nodes_to_numpy(bvh).view(oracle_mod.NODE_DTYPE),
                           bvh.prim_ids_tensor().cpu().numpy().view(np.uint32))


def _grad_tol(g64, scale, d):
    return 1e-5 * np.abs(g64) + 16 * EPS * scale / np.maximum(d, 1e-30)


@pytest.mark.parametrize("indexed", [True, False])
def test_gpu_fcp_triangles_bit_exact(cq, oracle_mod, indexed):
    print(f"[gpu] tris fcp indexed={indexed}", flush=True)
    from cubql_amd import synth
    V, F = synth.noisy_icosphere(4, seed=21)
    Q = synth.mesh_surface_queries(V, F, 20000, seed=22)
    tris = torch.from_numpy(F).cuda() if indexed else None
    verts = torch.from_numpy(V if indexed else V[F].reshape(-1, 3)).cuda()
    boxes = cq.triangles.boxes(tris, verts)
    assert np.array_equal(boxes.cpu().numpy(), oracle_mod.tri_boxes(F, V))
    bvh = cq.gpuBuilder(boxes, cq.BuildConfig(makeLeafThreshold=4))
    dQ = torch.from_numpy(Q).cuda()
    ids, d2, P = cq.triangles.findClosest(bvh, tris, verts, dQ)
    torch.cuda.synchronize()
    tree = _tree(cq, oracle_mod, bvh)
    oi, od, oP = oracle_mod.fcp_tris(tree, F if indexed else None,
                                     V if indexed else V[F].reshape(-1, 3), Q)
    assert np.array_equal(ids.cpu().numpy(), oi)
    assert np.array_equal(d2.cpu().numpy().view(np.uint32), od.view(np.uint32))
    assert np.array_equal(P.cpu().numpy(), oP)
    # FP64 error of the float32 distance (exact brute force on a sample)
    bi, bd = oracle_mod.fcp_tris_brute_f64(F, V, Q[:1500])
    dist32 = np.sqrt(od[:1500].astype(np.float64))
    dist64 = np.sqrt(bd)
    rel = np.abs(dist32 - dist64) / np.maximum(dist64, 1e-30)
    print(f"[gpu] tri distance rel err vs FP64: median {np.median(rel):.2e} "
          f"p99 {np.quantile(rel, 0.99):.2e} max {rel.max():.2e}", flush=True)
    assert np.all(np.abs(dist32 - dist64) <= 1e-5 * dist64 + 16 * EPS * 1.2)
    # a different id is only a tie: adjacent triangles sharing the closest
    # edge/vertex; its exact distance must equal the exact minimum
    for qi in np.nonzero(oi[:1500] != bi)[0]:
        _, dt = oracle_mod.fcp_tris_brute_f64(F[[oi[qi]]], V, Q[[qi]])
        assert abs(np.sqrt(dt[0]) - dist64[qi]) <= 1e-5 * dist64[qi] + 16 * EPS * 1.2
    bvh.free()


def test_gpu_triangle_backward(cq, oracle_mod):
    print("[gpu] tris backward", flush=True)
    from cubql_amd import synth
    V, F = synth.noisy_icosphere(4, seed=23)
    Q = synth.mesh_surface_queries(V, F, 20000, seed=24)
    dV, dF, dQ = torch.from_numpy(V).cuda(), torch.from_numpy(F).cuda(), torch.from_numpy(Q).cuda()
    bvh = cq.gpuBuilder(cq.triangles.boxes(dF, dV), cq.BuildConfig(makeLeafThreshold=4))
    ids, d2, _ = cq.triangles.findClosest(bvh, dF, dV, dQ, want_points=False)
    g = torch.linspace(0.5, 1.5, len(Q), device="cuda")
    gq = torch.empty_like(dQ)
    gv = torch.zeros_like(dV)
    cq.triangles.backward(dF, dV, dQ, ids, g, gq, gv)
    torch.cuda.synchronize()
    ids_h = ids.cpu().numpy()
    g64q, g64v = oracle_mod.tri_grad_f64(F, V, Q, ids_h, g.cpu().numpy().astype(np.float64))
    d = np.sqrt(d2.cpu().numpy().astype(np.float64))
    tol_q = _grad_tol(g64q, 1.3, d[:, None]) * 1.5 + 1e-7
    err_q = np.abs(gq.cpu().numpy() - g64q)
    print(f"[gpu] gradQ max abs err {err_q.max():.2e}; rel median "
          f"{np.median(err_q / np.maximum(np.abs(g64q), 1e-12)):.2e}", flush=True)
    assert np.all(err_q <= tol_q)
    # vertex gradients accumulate many contributions: bound by the sum of the
    # per-contribution tolerances
    contrib_tol = np.zeros_like(g64v)
    for j in range(3):
        np.add.at(contrib_tol, F[ids_h, j], np.abs(tol_q))
    err_v = np.abs(gv.cpu().numpy() - g64v)
    assert np.all(err_v <= contrib_tol + 1e-6)
    bvh.free()


@pytest.mark.parametrize("k", [1, 8])
def test_gpu_points_backward(cq, oracle_mod, k):
    print(f"[gpu] points backward k={k}", flush=True)
    from cubql_amd import synth
    P = synth.lcg_uniform(0, 100_000)
    Q = synth.lcg_uniform(1, 30_000)
    dP, dQ = torch.from_numpy(P).cuda(), torch.from_numpy(Q).cuda()
    bvh = cq.gpuBuilder(cq.points_to_boxes(dP), cq.BuildConfig(makeLeafThreshold=8))
    if k == 1:
        ids, d2 = cq.points.findClosest(bvh, dP, dQ)
        g = torch.ones(len(Q), device="cuda")
    else:
        ids, d2, _ = cq.points.findKNN(bvh, dP, dQ, k)
        g = torch.full((len(Q), k), 1.0 / (len(Q) * k), device="cuda")  # mean-of-kNN loss
    gq = torch.empty_like(dQ)
    gp = torch.zeros_like(dP)
    cq.points_backward(dP, dQ, ids, g, gq, gp)
    torch.cuda.synchronize()
    check_point_grads(oracle_mod, P, Q, ids.cpu().numpy(), g.cpu().numpy(), gq.cpu().numpy(),
                      gp.cpu().numpy(), label=f"C1 points k={k}")
    bvh.free()


def test_gpu_c3_workload(cq, oracle_mod):
    """C3 size: 1,310,720-triangle noisy icosphere, 4M surface queries,
    forward + backward; a 20K-query sample checked against the oracle."""
    print("[gpu] C3 full size", flush=True)
    from cubql_amd import synth
    V, F = synth.noisy_icosphere(8, seed=4)
    Q = synth.mesh_surface_queries(V, F, 4_000_000, seed=5)
    dV, dF, dQ = torch.from_numpy(V).cuda(), torch.from_numpy(F).cuda(), torch.from_numpy(Q).cuda()
    bvh = cq.gpuBuilder(cq.triangles.boxes(dF, dV), cq.BuildConfig(makeLeafThreshold=4))
    ids, d2, P = cq.triangles.findClosest(bvh, dF, dV, dQ)
    gq = torch.empty_like(dQ)
    gv = torch.zeros_like(dV)
    cq.triangles.backward(dF, dV, dQ, ids, torch.ones(len(Q), device="cuda"), gq, gv)
    torch.cuda.synchronize()
    tree = _tree(cq, oracle_mod, bvh)
    assert oracle_mod.check_tree(tree, oracle_mod.tri_boxes(F, V), 4)[0] == 0
    sel = np.arange(0, len(Q), len(Q) // 20000)
    oi, od, oP = oracle_mod.fcp_tris(tree, F, V, Q[sel])
    assert np.array_equal(ids.cpu().numpy()[sel], oi)
    assert np.array_equal(d2.cpu().numpy()[sel], od)
    g64q, _ = oracle_mod.tri_grad_f64(F, V, Q[sel], oi, np.ones(len(sel)))
    d = np.sqrt(od.astype(np.float64))[:, None]
    assert np.all(np.abs(gq.cpu().numpy()[sel] - g64q) <= _grad_tol(g64q, 1.3, d) * 1.5 + 1e-7)
    bvh.free()
