"""GPU parity of the batched path (cubql_batch_chamfer_f3).

* per cloud, the batched nearest neighbours and squared distances equal the
  plain per-cloud path (radix30Builder on the cloud + findClosest) bit for
  bit, in both directions;
* loss and gradients match an FP64 evaluation (oracle point_grad_f64) within
  the float32 tolerance of the adjoint (see tests/test_gpu_tris.py);
* results do not depend on the tunables (leaf size, block size, stack
  depth);
* a captured HIP graph replayed several times, with eager work between the
  replays, reproduces the eager result every time;
* edge cases: clouds with no source or no target points, single-point
  clouds, a batch of one cloud.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -24


@pytest.fixture(scope="module")
def cq():
    import cubql_amd
    cubql_amd.lib()
    return cubql_amd


def _batch(num, lo, hi, seed):
    from cubql_amd import synth
    ss = synth.chamfer_cloud_sizes(num, seed, lo, hi)
    ts = synth.chamfer_cloud_sizes(num, seed + 1, lo, hi)
    src, so = synth.chamfer_clouds_torch(ss, seed + 2, seed + 3)
    tgt, to = synth.chamfer_clouds_torch(ts, seed + 4, seed + 5)
    return src, so, tgt, to


def _plain(cq, src, so, tgt, to, T):
    """Per-cloud radix30Builder + findClosest, returned with global ids."""
    so_h, to_h = so.cpu().numpy(), to.cpu().numpy()
    out = []
    for A, ao, B, bo in [(src, so_h, tgt, to_h), (tgt, to_h, src, so_h)]:
        nn = np.full(A.shape[0], -1, np.int32)
        d2 = np.full(A.shape[0], np.inf, np.float32)
        for c in range(len(ao) - 1):
            a, b = A[ao[c]:ao[c + 1]], B[bo[c]:bo[c + 1]]
            if len(a) == 0 or len(b) == 0:
                continue
            bvh = cq.cuda.radix30Builder(cq.points_to_boxes(b), cq.BuildConfig(makeLeafThreshold=T))
            ids, dd = cq.points.findClosest(bvh, b, a)
            nn[ao[c]:ao[c + 1]] = ids.cpu().numpy() + bo[c]
            d2[ao[c]:ao[c + 1]] = dd.cpu().numpy()
            bvh.free()
        out.append((nn, d2))
    return out


def test_batch_equals_per_cloud_path(cq, oracle_mod):
    print("[gpu] batch vs per-cloud", flush=True)
    src, so, tgt, to = _batch(48, 64, 3000, 11)
    B = cq.ChamferBatch(src, so, tgt, to, cq.BatchParams(leafSize=8)).run()
    torch.cuda.synchronize()
    (pn_s, pd_s), (pn_t, pd_t) = _plain(cq, src, so, tgt, to, 8)
    assert np.array_equal(B.src_nn.cpu().numpy(), pn_s)
    assert np.array_equal(B.src_d2.cpu().numpy(), pd_s)
    assert np.array_equal(B.tgt_nn.cpu().numpy(), pn_t)
    assert np.array_equal(B.tgt_d2.cpu().numpy(), pd_t)
    # loss and gradients against FP64
    S, Tg = src.cpu().numpy(), tgt.cpu().numpy()
    so_h, to_h = so.cpu().numpy(), to.cpu().numpy()
    cs = np.repeat(np.arange(48), np.diff(so_h))
    ct = np.repeat(np.arange(48), np.diff(to_h))
    gs = 1.0 / np.diff(so_h)[cs]
    gt = 1.0 / np.diff(to_h)[ct]
    d_s = np.linalg.norm(S.astype(np.float64) - Tg[pn_s].astype(np.float64), axis=1)
    d_t = np.linalg.norm(Tg.astype(np.float64) - S[pn_t].astype(np.float64), axis=1)
    loss64 = np.bincount(cs, gs * d_s, 48) + np.bincount(ct, gt * d_t, 48)
    assert np.allclose(B.loss.cpu().numpy(), loss64, rtol=1e-5, atol=1e-6)
    gq_s, gp_t = oracle_mod.point_grad_f64(Tg, S, pn_s, gs)
    gq_t, gp_s = oracle_mod.point_grad_f64(S, Tg, pn_t, gt)
    g64_src, g64_tgt = gq_s + gp_s, gq_t + gp_t
    tol_s = 1e-5 * np.abs(g64_src) + 64 * EPS * (np.abs(g64_src).max() + 1e-3)
    assert np.all(np.abs(B.grad_src.cpu().numpy() - g64_src) <= tol_s + 1e-7)
    tol_t = 1e-5 * np.abs(g64_tgt) + 64 * EPS * (np.abs(g64_tgt).max() + 1e-3)
    assert np.all(np.abs(B.grad_tgt.cpu().numpy() - g64_tgt) <= tol_t + 1e-7)


@pytest.mark.parametrize("prm", [(2, 64, 32), (4, 128, 64), (16, 256, 32), (8, 256, 64)])
def test_batch_parameter_invariance(cq, prm):
    print(f"[gpu] batch params {prm}", flush=True)
    src, so, tgt, to = _batch(64, 64, 5000, 21)
    base = cq.ChamferBatch(src, so, tgt, to, cq.BatchParams(8, 256, 64)).run()
    other = cq.ChamferBatch(src, so, tgt, to, cq.BatchParams(*prm)).run()
    torch.cuda.synchronize()
    for a, b in [(base.src_nn, other.src_nn), (base.src_d2, other.src_d2),
                 (base.tgt_nn, other.tgt_nn), (base.tgt_d2, other.tgt_d2)]:
        assert torch.equal(a, b)


def test_batch_graph_replay(cq):
    print("[gpu] batch graph replay", flush=True)
    src, so, tgt, to = _batch(96, 64, 4000, 31)
    eager = cq.ChamferBatch(src, so, tgt, to).run()
    torch.cuda.synchronize()
    ref = [t.clone() for t in (eager.src_nn, eager.src_d2, eager.tgt_nn, eager.tgt_d2, eager.loss,
                               eager.grad_src, eager.grad_tgt)]
    g = cq.ChamferBatch(src, so, tgt, to).capture()
    for rep in range(4):
        for t in (g.src_nn, g.tgt_nn):
            t.fill_(-7)
        g.replay()
        torch.cuda.synchronize()
        # eager work between replays
        eager.run()
        torch.cuda.synchronize()
        assert torch.equal(g.src_nn, ref[0]) and torch.equal(g.src_d2, ref[1]), rep
        assert torch.equal(g.tgt_nn, ref[2]) and torch.equal(g.tgt_d2, ref[3]), rep
        assert torch.allclose(g.loss, ref[4], rtol=1e-5, atol=1e-6)
        assert torch.allclose(g.grad_src, ref[5], rtol=1e-4, atol=1e-6)
        assert torch.allclose(g.grad_tgt, ref[6], rtol=1e-4, atol=1e-6)


def test_batch_edge_cases(cq, oracle_mod):
    print("[gpu] batch edge cases", flush=True)
    rng = np.random.default_rng(4)
    ss = np.array([5, 0, 1, 300, 7, 1], dtype=np.int64)
    ts = np.array([0, 9, 1, 2, 700, 1], dtype=np.int64)
    S = torch.from_numpy(rng.random((ss.sum(), 3), dtype=np.float32)).cuda()
    T = torch.from_numpy(rng.random((ts.sum(), 3), dtype=np.float32)).cuda()
    so = torch.from_numpy(np.concatenate([[0], np.cumsum(ss)]).astype(np.int32)).cuda()
    to = torch.from_numpy(np.concatenate([[0], np.cumsum(ts)]).astype(np.int32)).cuda()
    B = cq.ChamferBatch(S, so, T, to, cq.BatchParams(leafSize=1)).run()
    torch.cuda.synchronize()
    nn = B.src_nn.cpu().numpy()
    Sh, Th = S.cpu().numpy(), T.cpu().numpy()
    so_h, to_h = so.cpu().numpy(), to.cpu().numpy()
    for c in range(len(ss)):
        a, b = Sh[so_h[c]:so_h[c + 1]], Th[to_h[c]:to_h[c + 1]]
        got = nn[so_h[c]:so_h[c + 1]]
        if len(b) == 0:
            assert np.all(got == -1)
        elif len(a):
            bi, _ = oracle_mod.fcp_brute(b, a)
            assert np.array_equal(got, bi + to_h[c])
    # one cloud
    B1 = cq.ChamferBatch(S[:5], so[:2], T[:9], torch.tensor([0, 9], dtype=torch.int32, device="cuda")).run()
    torch.cuda.synchronize()
    bi, _ = oracle_mod.fcp_brute(Th[:9], Sh[:5])
    assert np.array_equal(B1.src_nn.cpu().numpy(), bi)
