"""GPU parity of the fixed-box queries (cubql_box_query_f3 and the wide
variant): ids in traversal order and counts equal the oracle's
(orc_box_query, pinned in tests/test_oracle_box.py by brute-force sets and
closed-form orders), on spatial-median, Morton and SAH trees and their
4- and 8-wide collapses, with truncated outputs and count-only calls;
empty and degenerate queries (a point, an inverted box) included."""
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


@pytest.mark.parametrize("builder", ["sm", "radix30", "sah"])
def test_box_query_equals_oracle(cq, oracle_mod, builder):
    print(f"[gpu] box query {builder}", flush=True)
    rng = np.random.default_rng(5)
    P = oracle_mod.lcg_uniform(13, 200_000)
    B = np.concatenate([P - 0.001, P + 0.001], axis=1).astype(np.float32)
    c = rng.random((20_000, 3)).astype(np.float32)
    h = (rng.random((20_000, 3)) * 0.02).astype(np.float32)
    Q = np.concatenate([c - h, c + h], axis=1).astype(np.float32)
    Q[0] = [0.5, 0.5, 0.5, 0.5, 0.5, 0.5]          # a point
    Q[1] = [0.6, 0.6, 0.6, 0.4, 0.4, 0.4]          # inverted: overlaps nothing
    Q[2] = [-1, -1, -1, 2, 2, 2]                   # everything
    dB, dQ = torch.from_numpy(B).cuda(), torch.from_numpy(Q).cuda()
    cfg = cq.BuildConfig(makeLeafThreshold=4)
    bvh = {"sm": cq.gpuBuilder, "radix30": cq.cuda.radix30Builder, "sah": cq.cuda.sahBuilder}[builder](dB, cfg)
    torch.cuda.synchronize()
    tree = _tree(cq, oracle_mod, bvh)
    _, cnt = cq.fixedBoxQuery.forEachPrim(bvh, dB, dQ)
    torch.cuda.synchronize()
    ocnt = oracle_mod.box_query(tree, B, Q, 0)[1]
    assert np.array_equal(cnt.cpu().numpy().astype(np.uint32), ocnt)
    assert ocnt[1] == 0 and ocnt[2] == len(P)
    m = int(np.sort(ocnt)[-4])  # truncates the three largest lists
    ids, cnt2 = cq.fixedBoxQuery.forEachPrim(bvh, dB, dQ, m)
    torch.cuda.synchronize()
    oids, _ = oracle_mod.box_query(tree, B, Q, m)
    got = ids.cpu().numpy().view(np.uint32)
    assert np.array_equal(cnt2.cpu().numpy().astype(np.uint32), ocnt)
    for q in range(len(Q)):
        k = min(m, int(ocnt[q]))
        assert np.array_equal(got[q, :k], oids[q, :k]), q
    for W in (4, 8):
        wide = cq.collapse_wide(bvh, W)
        wids, wcnt = cq.fixedBoxQuery.forEachPrim(wide, dB, dQ, m)
        torch.cuda.synchronize()
        wref = oracle_mod.wide_collapse(tree, W)
        oids, ocw = oracle_mod.box_query(tree, B, Q, m, wide=wref, width=W)
        assert np.array_equal(wcnt.cpu().numpy().astype(np.uint32), ocw)
        g = wids.cpu().numpy().view(np.uint32)
        for q in range(len(Q)):
            k = min(m, int(ocw[q]))
            assert np.array_equal(g[q, :k], oids[q, :k]), (W, q)
        wide.free()
    bvh.free()
