"""GPU parity of the batched path (cubql_batch_chamfer_f3).

* per cloud, the batched nearest neighbours and squared distances equal the
  plain per-cloud path (radix30Builder on the cloud + findClosest) bit for
  bit, in both directions;
* loss and gradients match an FP64 evaluation (oracle point_grad_f64) within
  1e-5 of each point's own gradient size (rule in tests/gradtol.py);
* the full uneven batch (4096 cloud pairs, sizes log-uniform in [64, 65536],
  about 38M points per side, the benchmark's seeds): nearest ids and squared
  distances of 64+ clouds (the 10 largest, the 10 smallest, random ones)
  equal the oracle's, loss and gradients match FP64, the graph replay
  equals the eager run and the tunables do not change the result;
* the stack-overflow path: with an 8- or 16-entry fast stack many queries
  rerun in the deep kernel (clouds of thousands of identical points among
  them), and still equal the 64-entry run and brute force;
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
from gradtol import check_chamfer_grads

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
    # loss and gradients against FP64, cloud by cloud (tests/gradtol.py)
    _check_loss_and_grads(oracle_mod, B, src, so, tgt, to, range(48))


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
    # offsets that do not split a side into clouds are refused on the host
    # (no kernel runs)
    bad = [torch.tensor(o, dtype=torch.int32, device="cuda") for o in ([1, 5], [0, 4], [0, 6, 5])]
    for o in bad:
        with pytest.raises(cq.CubqlError):
            cq.ChamferBatch(S[:5], o, T[:9], torch.tensor([0, 9] + [9] * (o.numel() - 2), dtype=torch.int32,
                                                          device="cuda"))


def _check_loss_and_grads(oracle_mod, B, src, so, tgt, to, clouds):
    S, Tg = src.cpu().numpy(), tgt.cpu().numpy()
    so_h, to_h = so.cpu().numpy().astype(np.int64), to.cpu().numpy().astype(np.int64)
    nn_s, nn_t = B.src_nn.cpu().numpy(), B.tgt_nn.cpu().numpy()
    loss = B.loss.cpu().numpy()
    gs, gt = B.grad_src.cpu().numpy(), B.grad_tgt.cpu().numpy()
    worst = 0.0
    for c in clouds:
        sb, se, tb, te = so_h[c], so_h[c + 1], to_h[c], to_h[c + 1]
        if se == sb or te == tb:
            continue
        a, b = S[sb:se], Tg[tb:te]
        ls, lt = nn_s[sb:se] - tb, nn_t[tb:te] - sb
        d_s = np.linalg.norm(a.astype(np.float64) - b[ls].astype(np.float64), axis=1)
        d_t = np.linalg.norm(b.astype(np.float64) - a[lt].astype(np.float64), axis=1)
        l64 = d_s.mean() + d_t.mean()
        assert abs(loss[c] - l64) <= 1e-5 * l64 + 1e-30, (c, loss[c], l64)
        worst = max(worst, check_chamfer_grads(oracle_mod, a, b, ls, lt, gs[sb:se], gt[tb:te],
                                               label=f"cloud {c}"))
    print(f"[gpu] chamfer gradient rel err (of each point's sum|g|): max {worst:.2e}", flush=True)


def _oracle_nn(oracle_mod, data, queries):
    """Nearest data point per query, lowest index on ties: brute force for
    small clouds, the oracle's tree query (same rule) for large ones."""
    if len(data) * len(queries) <= 4e7:
        return oracle_mod.fcp_brute(data, queries)
    tree = oracle_mod.radix_build(oracle_mod.points_to_boxes(data), 10, 8)
    return oracle_mod.fcp(tree, data, queries, lowest_id=True)


def test_batch_c5_full_config(cq, oracle_mod):
    """4096 uneven cloud pairs at the benchmark's seeds and sizes."""
    print("[gpu] batch C5 full config", flush=True)
    from cubql_amd import synth
    ss = synth.chamfer_cloud_sizes(4096, 101)
    ts = synth.chamfer_cloud_sizes(4096, 102)
    src, so = synth.chamfer_clouds_torch(ss, 103, 104)
    tgt, to = synth.chamfer_clouds_torch(ts, 105, 106)
    maxc = int(max(ss.max(), ts.max()))
    print(f"[gpu] C5: {len(src):,} source / {len(tgt):,} target points, largest cloud {maxc}",
          flush=True)
    B = cq.ChamferBatch(src, so, tgt, to, cq.BatchParams(8, 128, 32, maxc)).run()
    G = cq.ChamferBatch(src, so, tgt, to, cq.BatchParams(4, 256, 64, maxc)).capture()
    for _ in range(2):
        G.replay()
    torch.cuda.synchronize()
    for a, b in [(B.src_nn, G.src_nn), (B.src_d2, G.src_d2), (B.tgt_nn, G.tgt_nn),
                 (B.tgt_d2, G.tgt_d2)]:
        assert torch.equal(a, b)
    tot = ss + ts
    order = np.argsort(tot, kind="stable")
    rng = np.random.default_rng(7)
    sel = set(order[-10:].tolist()) | set(order[:10].tolist())
    sel |= set(int(x) for x in np.argsort(ss)[-3:]) | set(int(x) for x in np.argsort(ts)[-3:])
    sel |= set(rng.choice(4096, 44, replace=False).tolist())
    sel = sorted(sel)
    assert len(sel) >= 64
    S, Tg = src.cpu().numpy(), tgt.cpu().numpy()
    so_h, to_h = so.cpu().numpy().astype(np.int64), to.cpu().numpy().astype(np.int64)
    nn_s, d_s = B.src_nn.cpu().numpy(), B.src_d2.cpu().numpy()
    nn_t, d_t = B.tgt_nn.cpu().numpy(), B.tgt_d2.cpu().numpy()
    for c in sel:
        a, b = S[so_h[c]:so_h[c + 1]], Tg[to_h[c]:to_h[c + 1]]
        oi, od = _oracle_nn(oracle_mod, b, a)
        assert np.array_equal(nn_s[so_h[c]:so_h[c + 1]], oi + to_h[c]), f"cloud {c} src->tgt ids"
        assert np.array_equal(d_s[so_h[c]:so_h[c + 1]], od), f"cloud {c} src->tgt distances"
        oi, od = _oracle_nn(oracle_mod, a, b)
        assert np.array_equal(nn_t[to_h[c]:to_h[c + 1]], oi + so_h[c]), f"cloud {c} tgt->src ids"
        assert np.array_equal(d_t[to_h[c]:to_h[c + 1]], od), f"cloud {c} tgt->src distances"
    print(f"[gpu] C5: {len(sel)} clouds equal the oracle (largest {int(tot[order[-1]])} points)",
          flush=True)
    _check_loss_and_grads(oracle_mod, B, src, so, tgt, to, sel)


@pytest.mark.parametrize("stack", [8, 16])
def test_batch_stack_overflow_path(cq, oracle_mod, stack):
    """A shallow fast-path stack (8 or 16 entries) makes many queries
    overflow, so the deep rerun carries a large share of them; results must
    equal the 64-entry run and brute force.  Clouds of many identical points
    (both children of every node at the same distance, one far child stacked
    per level) are included."""
    print(f"[gpu] batch stack overflow path stack={stack}", flush=True)
    src, so, tgt, to = _batch(24, 64, 4000, 41)
    rng = np.random.default_rng(12)
    S, T = [src.cpu().numpy()], [tgt.cpu().numpy()]
    so_l, to_l = list(so.cpu().numpy()), list(to.cpu().numpy())
    for n in [3000, 70, 5000]:
        dup = np.tile(rng.random((1, 3), dtype=np.float32) * 10, (n, 1))
        S.append(np.concatenate([dup, rng.random((5, 3), dtype=np.float32) * 10]))
        T.append(np.concatenate([np.tile(dup[:1], (n // 2 + 1, 1)),
                                 rng.random((n // 3 + 1, 3), dtype=np.float32) * 10]))
        so_l.append(so_l[-1] + len(S[-1]))
        to_l.append(to_l[-1] + len(T[-1]))
    S, T = np.concatenate(S), np.concatenate(T)
    so_h, to_h = np.asarray(so_l, np.int32), np.asarray(to_l, np.int32)
    dS, dT = torch.from_numpy(S).cuda(), torch.from_numpy(T).cuda()
    dso, dto = torch.from_numpy(so_h).cuda(), torch.from_numpy(to_h).cuda()
    B = cq.ChamferBatch(dS, dso, dT, dto, cq.BatchParams(1, 128, stack)).run()
    R = cq.ChamferBatch(dS, dso, dT, dto, cq.BatchParams(1, 128, 64)).run()
    torch.cuda.synchronize()
    for a, b in [(B.src_nn, R.src_nn), (B.src_d2, R.src_d2), (B.tgt_nn, R.tgt_nn),
                 (B.tgt_d2, R.tgt_d2)]:
        assert torch.equal(a, b)
    nn_s, nn_t, d_s = B.src_nn.cpu().numpy(), B.tgt_nn.cpu().numpy(), B.src_d2.cpu().numpy()
    for c in range(len(so_h) - 1):
        a, b = S[so_h[c]:so_h[c + 1]], T[to_h[c]:to_h[c + 1]]
        bi, bd = oracle_mod.fcp_brute(b, a)
        assert np.array_equal(nn_s[so_h[c]:so_h[c + 1]], bi + to_h[c]), c
        assert np.array_equal(d_s[so_h[c]:so_h[c + 1]], bd), c
        bi, bd = oracle_mod.fcp_brute(a, b)
        assert np.array_equal(nn_t[to_h[c]:to_h[c + 1]], bi + so_h[c]), c
    _check_loss_and_grads(oracle_mod, B, dS, dso, dT, dto, range(len(so_h) - 1))


@pytest.mark.parametrize("stack", [32, 64])
def test_batch_stack_overflow_clouds(cq, oracle_mod, stack):
    """Clouds of many identical points: every node on a query's path has
    both children at the same distance, so the traversal stacks one far
    child per level and overflows any stack shallower than the tree (30
    Morton levels are equal, then the equal-key middle splits); those
    queries rerun in the deep kernel and must still equal brute force."""
    print(f"[gpu] batch stack overflow clouds stack={stack}", flush=True)
    rng = np.random.default_rng(12)
    clouds_s, clouds_t = [], []
    for c, n in enumerate([3000, 70, 5000, 1200]):
        dup = np.tile(rng.random((1, 3), dtype=np.float32) * 10, (n, 1))
        clouds_s.append(np.concatenate([dup, rng.random((5, 3), dtype=np.float32) * 10]))
        clouds_t.append(np.concatenate([np.tile(dup[:1], (n // 2 + 1, 1)),
                                        rng.random((n // 3 + 1, 3), dtype=np.float32) * 10]))
    S = np.concatenate(clouds_s)
    T = np.concatenate(clouds_t)
    so = np.concatenate([[0], np.cumsum([len(x) for x in clouds_s])]).astype(np.int32)
    to = np.concatenate([[0], np.cumsum([len(x) for x in clouds_t])]).astype(np.int32)
    dS, dT = torch.from_numpy(S).cuda(), torch.from_numpy(T).cuda()
    B = cq.ChamferBatch(dS, torch.from_numpy(so).cuda(), dT, torch.from_numpy(to).cuda(),
                        cq.BatchParams(1, 128, stack)).run()
    torch.cuda.synchronize()
    nn_s, nn_t = B.src_nn.cpu().numpy(), B.tgt_nn.cpu().numpy()
    for c in range(len(clouds_s)):
        a, b = S[so[c]:so[c + 1]], T[to[c]:to[c + 1]]
        bi, bd = oracle_mod.fcp_brute(b, a)
        assert np.array_equal(nn_s[so[c]:so[c + 1]], bi + to[c]), c
        assert np.array_equal(B.src_d2.cpu().numpy()[so[c]:so[c + 1]], bd), c
        bi, bd = oracle_mod.fcp_brute(a, b)
        assert np.array_equal(nn_t[to[c]:to[c + 1]], bi + so[c]), c
    _check_loss_and_grads(oracle_mod, B, dS, torch.from_numpy(so), dT, torch.from_numpy(to),
                          range(len(clouds_s)))
