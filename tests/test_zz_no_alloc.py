"""cubql_fcp_points_noalloc_f3 (the closest-point entry point that allocates
nothing) against cubql_fcp_points_f3, through the C ABI.

Two calls in a row on the shared static workspace, the second one smaller
and over another tree: ids and squared distances must equal the allocating
entry point's bit for bit each time.  The first call is large enough to take
the query-ordering path (which uses most of the workspace); the second is
below that size, so stale workspace contents from the first call must not
leak into it.  A call over the query limit is refused on the host.
"""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

MAX_Q = 1048576
ERR_INVALID_ARG = -1


@pytest.fixture(scope="module")
def cq():
    import cubql_amd
    cubql_amd.lib()
    return cubql_amd


def _both(cq, bvh, P, Q):
    L = cq.lib()
    nq = Q.shape[0]
    out = []
    for fn in (L.cubql_fcp_points_f3, L.cubql_fcp_points_noalloc_f3):
        ids = torch.full((nq,), -7, dtype=torch.int32, device="cuda")
        d2 = torch.full((nq,), -7.0, dtype=torch.float32, device="cuda")
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        rc = fn(C.byref(bvh.c), C.c_void_p(P.data_ptr()), C.c_void_p(Q.data_ptr()), nq, C.c_float(np.inf),
                C.c_void_p(ids.data_ptr()), C.c_void_p(d2.data_ptr()), st)
        assert rc == 0, (fn.__name__, rc, L.cubql_last_error_string())
        torch.cuda.synchronize()
        out.append((ids.cpu().numpy(), d2.cpu().numpy()))
    return out


def test_noalloc_fcp_equals_allocating_entry_point(cq, oracle_mod):
    from cubql_amd import synth
    calls = [
        # (points, queries, seed of points, seed of queries)
        (300_000, 400_000, 0, 1),   # query-ordered path (>= 32768 queries)
        (20_000, 5_000, 2, 3),      # smaller: input order, stale workspace
    ]
    for i, (n, nq, sp, sq) in enumerate(calls):
        print(f"[gpu] noalloc call {i}: {n} points, {nq} queries", flush=True)
        P = torch.from_numpy(synth.lcg_uniform(sp, n)).cuda()
        Q = torch.from_numpy(synth.lcg_uniform(sq, nq, -0.1, 1.1)).cuda()
        bvh = cq.gpuBuilder(cq.points_to_boxes(P), cq.BuildConfig(makeLeafThreshold=8))
        (ia, da), (ib, db) = _both(cq, bvh, P, Q)
        assert np.array_equal(ia, ib), f"call {i}: ids differ"
        assert np.array_equal(da, db), f"call {i}: distances differ"
        assert np.all(ib >= 0) and np.all(ib < n)
        if i == 1:
            # small enough for the oracle's brute force
            bi, bd = oracle_mod.fcp_brute(P.cpu().numpy(), Q.cpu().numpy())
            assert np.array_equal(ib, bi) and np.array_equal(db, bd)
        bvh.free()


def test_noalloc_fcp_rejects_too_many_queries(cq):
    L = cq.lib()
    P = torch.zeros((8, 3), dtype=torch.float32, device="cuda")
    bvh = cq.gpuBuilder(cq.points_to_boxes(P), cq.BuildConfig())
    # rejected before any pointer is used: the output buffers stay unread
    nq = MAX_Q + 1
    Q = torch.zeros((16, 3), dtype=torch.float32, device="cuda")
    rc = L.cubql_fcp_points_noalloc_f3(C.byref(bvh.c), C.c_void_p(P.data_ptr()), C.c_void_p(Q.data_ptr()), nq,
                                       C.c_float(np.inf), C.c_void_p(Q.data_ptr()), None,
                                       C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == ERR_INVALID_ARG
    assert b"limit" in L.cubql_last_error_string()
    bvh.free()
