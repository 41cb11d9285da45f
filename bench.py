#!/usr/bin/env python3
"""bench.py -- BVH build-and-query benchmark on one MI355X.

Workload (the large GPU configuration the headline metric is quoted on):
10,000,000 float32 points uniform in [0,1]^3 (reference LCG<8> generator,
seed 2), each a degenerate box; gpuBuilder spatial median with leaf size <= 8.
A step is one complete build (allocation and free included, as the
reference's buildBench does, testing/bvh/buildBench.h:115-143) with the
boxes already resident in HBM.  After the timed steps the same run measures
10M closest-point queries (seed 3) and 1M kNN queries (k=8) on the built
tree, the per-kernel roofline of the dominant build kernel (HIP events
inside the library, on the build stream), and the reference's own CPU path
on a bounded sample.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0           # MI355X HBM3E
PUBLISHED_MPRIMS = 1e7 / 13e-3 / 1e6  # "10M points in under 13ms" (README.md:107-110)


def cpu_info():
    cores = os.cpu_count()
    model = platform.processor() or ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except OSError:
        pass
    return cores, model


def _ref_method(fn, min_reps, min_s):
    """The reference's timing method (testing/bvh/buildBench.h:115-133): one
    warm-up, then repeat until at least `min_reps` runs and `min_s` seconds."""
    fn()
    times = []
    t0 = time.perf_counter()
    while len(times) < min_reps or time.perf_counter() - t0 < min_s:
        times.append(fn())
    return times


def cpu_baseline(n_sample: int, nq_sample: int, min_s: float):
    """The reference's own CPU path (cpu::spatialMedian + host findClosest /
    findKNN, compiled from its sources into oracle/_ref), 1 core, on bounded
    samples of the benchmark workloads; plus the same per-query function on
    all host cores from a thread pool of ours (ctypes drops the GIL)."""
    import concurrent.futures as cf

    import numpy as np

    import oracle
    from cubql_amd import synth

    kind = "reference" if oracle.ref_available() else "port"

    def timed_build(boxes, n):
        if kind == "reference":
            return lambda: oracle.ref().ref_time_cpu_build(oracle._p(boxes), n, 8)

        def once():
            t = time.perf_counter()
            oracle.sm_build(boxes, 8)
            return time.perf_counter() - t
        return once

    def tree_of(boxes):
        return oracle.ref_sm_build(boxes, 8) if kind == "reference" else oracle.sm_build(boxes, 8)

    def fcp(tree, P, Q):
        return oracle.ref_fcp(tree, P, Q) if kind == "reference" else oracle.fcp(tree, P, Q, lowest_id=False)

    def knn(tree, P, Q, k):
        return oracle.ref_knn(tree, P, Q, k) if kind == "reference" else oracle.knn(tree, P, Q, k)

    def rate(fn, units):
        t = time.perf_counter()
        fn()
        return units / (time.perf_counter() - t) / 1e6

    # 10M uniform points (seed 2): the first n_sample of the same stream
    P = synth.lcg_uniform(2, n_sample)
    Q = synth.lcg_uniform(3, nq_sample)
    boxes = oracle.points_to_boxes(P)
    times = _ref_method(timed_build(boxes, n_sample), 10, min_s)
    tree = tree_of(boxes)
    fcp_rate = rate(lambda: fcp(tree, P, Q), nq_sample)
    nk = nq_sample // 4
    knn_rate = rate(lambda: knn(tree, P, Q[:nk], 8), nk)
    # all host cores: our thread pool over the unchanged per-query function
    threads = min(os.cpu_count() or 1, 16)
    Qa = synth.lcg_uniform(3, nq_sample * 4)
    chunks = np.array_split(Qa, threads * 4)
    t = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda q: fcp(tree, P, q), chunks))
    fcp_all = len(Qa) / (time.perf_counter() - t) / 1e6

    # 100K points (seed 0), 100K queries in [-0.1, 1.1]^3 (seed 1): full size
    P1 = synth.lcg_uniform(0, 100_000)
    Q1 = synth.lcg_uniform(1, 100_000, -0.1, 1.1)
    b1 = oracle.points_to_boxes(P1)
    t1 = _ref_method(timed_build(b1, len(P1)), 10, min(min_s, 3.0))
    tree1 = tree_of(b1)
    c1_fcp = rate(lambda: fcp(tree1, P1, Q1), len(Q1))

    # 2M-point Gaussian mixture: spatial median (the reference has no CPU
    # Morton builder) + findKNN k=16 over a query subset
    P4 = synth.gaussian_mixture(2_000_000, seed=41)
    Q4 = synth.gaussian_mixture(2_000_000, seed=42)[: nq_sample // 4]
    b4 = oracle.points_to_boxes(P4)
    t4 = _ref_method(timed_build(b4, len(P4)), 3, min(min_s, 3.0))
    tree4 = tree_of(b4)
    c4_knn = rate(lambda: knn(tree4, P4, Q4, 16), len(Q4))

    cores, model = cpu_info()
    mean = float(np.mean(times))
    return {
        "value": n_sample / mean / 1e6,
        "unit": "Mprims/s",
        "cores": 1,
        "kind": kind,
        "sample": (f"cpu::spatialMedian (serial, leaf<=8) over the first {n_sample:,} points of the "
                   f"10M seed-2 stream, {len(times)} timed builds after 1 warm-up, "
                   f"{sum(times):.1f} s (mean {mean*1e3:.1f} ms, median "
                   f"{float(np.median(times))*1e3:.1f} ms); host findClosest over {nq_sample:,} and "
                   f"findKNN k=8 over {nk:,} seed-3 queries on that tree"),
        "method": "1 warm-up, then >= 10 runs and >= %.0f s (testing/bvh/buildBench.h:115-133)" % min_s,
        "fcp_mqueries_s": fcp_rate,
        "knn8_mqueries_s": knn_rate,
        "fcp_mqueries_s_all_cores": fcp_all,
        "all_cores_threads": threads,
        "all_cores_note": (f"the reference's host findClosest over {len(Qa):,} queries from a "
                           f"{threads}-thread pool of ours"),
        "c1": {"build_ms_mean": float(np.mean(t1)) * 1e3, "build_ms_median": float(np.median(t1)) * 1e3,
               "builds": len(t1), "fcp_mqueries_s": c1_fcp,
               "sample": ("full size: 100,000 points (seed 0), 100,000 queries (seed 1); >= 10 builds, "
                          ">= 3 s")},
        "c4": {"build_ms_mean": float(np.mean(t4)) * 1e3, "builds": len(t4), "knn16_mqueries_s": c4_knn,
               "sample": (f"cpu::spatialMedian over the full 2M-point mixture (>= 3 builds, >= 3 s); "
                          f"findKNN k=16 over "
                          f"{len(Q4):,} mixture queries")},
        "compiler": "g++ -O3 -std=c++17 -ffp-contract=off (x86-64 default -march)",
        "host_cpus": cores,
        "host_model": model,
    }


def bench_batched(cq, dev, num_clouds: int, plain_budget_s: float, tune: bool):
    """C5: batched uneven Chamfer fwd+bwd vs the plain per-cloud path."""
    import numpy as np
    import torch

    from cubql_amd import synth
    from cubql_amd.autotune import tune_batch

    ss = synth.chamfer_cloud_sizes(num_clouds, 101)
    ts = synth.chamfer_cloud_sizes(num_clouds, 102)
    src, so = synth.chamfer_clouds_torch(ss, 103, 104, device=dev)
    tgt, to = synth.chamfer_clouds_torch(ts, 105, 106, device=dev)
    torch.cuda.synchronize()
    maxc = int(max(ss.max(), ts.max()))
    if tune:
        # stack depths 8 and 16 run the traversal stack in LDS, 64 in scratch
        cands = [(T, b, st) for T in (4, 8, 16) for b in (128, 256) for st in (8, 16, 64)]
        params, times = tune_batch(src, so, tgt, to, cands, reps=2, max_cloud_size=maxc)
    else:
        params, times = cq.BatchParams(8, 256, 64, maxc), {}
    batch = cq.ChamferBatch(src, so, tgt, to, params).capture()
    for _ in range(2):
        batch.replay()
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.replay()
    torch.cuda.synchronize()
    batched_ms = (time.perf_counter() - t0) / reps * 1e3
    # plain: per cloud, radix30Builder on both sides + fcp both ways + adjoint
    so_h, to_h = so.cpu().numpy(), to.cpu().numpy()
    mem = cq.TorchMemory(dev)
    cfg = cq.BuildConfig(makeLeafThreshold=params.leafSize)
    gs = torch.empty_like(src)
    gt = torch.empty_like(tgt)
    done = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for c in range(num_clouds):
        S, T = src[so_h[c]:so_h[c + 1]], tgt[to_h[c]:to_h[c + 1]]
        bt = cq.cuda.radix30Builder(cq.points_to_boxes(T), cfg, mem=mem)
        bs = cq.cuda.radix30Builder(cq.points_to_boxes(S), cfg, mem=mem)
        nn_s, d_s = cq.points.findClosest(bt, T, S)
        nn_t, d_t = cq.points.findClosest(bs, S, T)
        g_s = torch.full((S.shape[0],), 1.0 / S.shape[0], device=dev)
        g_t = torch.full((T.shape[0],), 1.0 / T.shape[0], device=dev)
        cq.points_backward(T, S, nn_s, g_s, gs[so_h[c]:so_h[c + 1]], gt[to_h[c]:to_h[c + 1]])
        cq.points_backward(S, T, nn_t, g_t, gt[to_h[c]:to_h[c + 1]], gs[so_h[c]:so_h[c + 1]])
        bt.free()
        bs.free()
        done += 1
        if time.perf_counter() - t0 > plain_budget_s:
            break
    torch.cuda.synchronize()
    plain_ms_done = (time.perf_counter() - t0) * 1e3
    frac_pts = (so_h[done] + to_h[done]) / (so_h[-1] + to_h[-1])
    plain_ms = plain_ms_done if done == num_clouds else plain_ms_done / max(done / num_clouds, 1e-9)
    return {
        "clouds": num_clouds, "src_points": int(so_h[-1]), "tgt_points": int(to_h[-1]),
        "params": {"leafSize": params.leafSize, "blockSize": params.blockSize,
                   "stackDepth": params.stackDepth},
        "tuning_ms": times,
        "batched_ms": batched_ms, "batched_mpoints_s": (so_h[-1] + to_h[-1]) / batched_ms / 1e3,
        "plain_ms": plain_ms, "plain_clouds_timed": done,
        "plain_note": ("all clouds timed" if done == num_clouds else
                       f"first {done} clouds timed ({frac_pts:.1%} of the points), scaled by cloud count"),
        "speedup_vs_plain": plain_ms / batched_ms,
        "batched_is_graph_replay": True,
    }


def partition_traffic():
    """HBM bytes per sm::k_pair launch (the pass-B partition, reported as
    sm_partition) measured with rocprofv3 PMC
    counters (FETCH_SIZE + WRITE_SIZE, scaled by a 1 GiB copy); the summary
    with its commands lives in profiles/sm_partition_10M_pmc.md."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "sm_partition_10M_pmc.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def _time_cuda(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _kernel_ms(cq, name, fn, reps=3):
    """Average duration of one launch of kernel `name`, from the library's
    HIP events on the launch stream (cubql_profile_*)."""
    import torch
    cq.profile.reset()
    cq.profile.enable(True)
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    cq.profile.enable(False)
    st = cq.profile.query(name)
    return st["ms"] / st["launches"] if st and st["launches"] else None


def query_roofline(cq, stats, kernel_ms, io_bytes, prim_bytes, kernel):
    """Visited-bytes roofline of a query kernel: per query, `io_bytes` of
    query in and result out, the node bytes the traversal loaded (64 per
    child pair expanded, 32 per node popped) and `prim_bytes` per prim
    tested, all counted by an instrumented run of the same queries."""
    nq = stats["queries"]
    total = nq * io_bytes + stats["node_bytes"] + stats["prims"] * prim_bytes
    gbs = total / (kernel_ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": kernel, "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": gbs / HBM_PEAK_GBS, "avg_launch_ms": kernel_ms, "bytes_per_launch": total,
            "bytes_per_query": total / nq, "node_bytes_per_query": stats["node_bytes"] / nq,
            "prims_per_query": stats["prims"] / nq,
            "bytes_model": f"{io_bytes} B query+result, node bytes loaded, {prim_bytes} B per prim tested",
            "note": "visited bytes, not HBM traffic: neighbouring (Morton-ordered) queries load the same "
                    "nodes, most of them from L2 or the MALL, so frac can pass 1"}


def bench_triangles(cq, dev, reps: int):
    """1,310,720-triangle noisy icosphere: SAH build (32 bins, leaves <= 4),
    4M closest-triangle queries forward and backward."""
    import torch
    from cubql_amd import synth
    V, F = synth.noisy_icosphere(8, seed=4)
    Q = synth.mesh_surface_queries(V, F, 4_000_000, seed=5)
    dV, dF, dQ = (torch.from_numpy(V).to(dev), torch.from_numpy(F).to(dev),
                  torch.from_numpy(Q).to(dev))
    mem = cq.TorchMemory(dev)
    cfg = cq.BuildConfig(makeLeafThreshold=4, maxAllowedLeafSize=4)
    boxes = cq.triangles.boxes(dF, dV)
    build_ms = _time_cuda(lambda: cq.cuda.sahBuilder(boxes, cfg, mem=mem, numBins=32).free(), reps)
    cq.profile.reset()
    cq.profile.enable(True)
    cq.cuda.sahBuilder(boxes, cfg, mem=mem, numBins=32).free()
    torch.cuda.synchronize()
    cq.profile.enable(False)
    kern = {}
    for name in ["sah_bin", "sah_partition", "sah_small"]:
        st = cq.profile.query(name)
        if st and st["launches"]:
            kern[name] = {"launches_per_build": st["launches"], "ms_per_build": st["ms"],
                          "gbs": st["bytes"] / (st["ms"] * 1e-3) / 1e9}
    bvh = cq.cuda.sahBuilder(boxes, cfg, mem=mem, numBins=32)
    out = {}

    def fwd():
        out["r"] = cq.triangles.findClosest(bvh, dF, dV, dQ, want_points=False)
    fwd_ms = _time_cuda(fwd, reps)
    kms = _kernel_ms(cq, "fcp_triangles", fwd)
    stats = cq.query_stats(bvh, dQ, tris=dF, verts=dV)
    # query 12 B in, id + distance 8 B out; per triangle 4 B primID,
    # 12 B index triple and 3 x 12 B vertices
    roof = query_roofline(cq, stats, kms, 20, 52, "fcp_triangles") if kms else None
    ids = out["r"][0]
    g = torch.ones(len(Q), device=dev)
    gq = torch.empty_like(dQ)
    gv = torch.zeros_like(dV)
    bwd_ms = _time_cuda(lambda: cq.triangles.backward(dF, dV, dQ, ids, g, gq, gv), reps)
    res = {"triangles": int(len(F)), "queries": int(len(Q)), "sah_bins": 32, "leaf": 4,
           "sah_build_ms": build_ms, "sah_build_mprims_s": len(F) / build_ms / 1e3,
           "num_nodes": int(bvh.numNodes),
           "fcp_ms": fwd_ms, "fcp_mqueries_s": len(Q) / fwd_ms / 1e3,
           "backward_ms": bwd_ms, "fwd_bwd_mqueries_s": len(Q) / (fwd_ms + bwd_ms) / 1e3,
           "sah_kernels": kern, "fcp_roofline": roof}
    bvh.free()
    return res


def bench_mixture(cq, dev, reps: int):
    """2M-point Gaussian mixture with 1% outliers at 1e6: rebinning Morton
    build (leaves <= 8), kNN k=16 over 2M queries, mean-kNN backward."""
    import torch
    from cubql_amd import synth
    P = torch.from_numpy(synth.gaussian_mixture(2_000_000, seed=41)).to(dev)
    Q = torch.from_numpy(synth.gaussian_mixture(2_000_000, seed=42)).to(dev)
    mem = cq.TorchMemory(dev)
    cfg = cq.BuildConfig(makeLeafThreshold=8)
    boxes = cq.points_to_boxes(P)
    build_ms = _time_cuda(lambda: cq.cuda.rebinRadixBuilder(boxes, cfg, mem=mem).free(), reps)
    bvh = cq.cuda.rebinRadixBuilder(boxes, cfg, mem=mem)
    out = {}

    def fwd():
        out["r"] = cq.points.findKNN(bvh, P, Q, 16)
    knn_ms = _time_cuda(fwd, reps)
    kms = _kernel_ms(cq, "knn_points", fwd)
    stats = cq.query_stats(bvh, Q, data=P, k=16)
    # query 12 B in, 16 x (id + distance) + count out; 16 B per point tested
    roof = query_roofline(cq, stats, kms, 12 + 8 * 16 + 4, 16, "knn_points<16>") if kms else None
    ids = out["r"][0]
    g = torch.full((len(Q), 16), 1.0 / (len(Q) * 16), device=dev)
    gq = torch.empty_like(Q)
    gp = torch.zeros_like(P)
    bwd_ms = _time_cuda(lambda: cq.points_backward(P, Q, ids, g, gq, gp), reps)
    # the same without the scattered point gradient: what the scatter costs
    bwd_q_ms = _time_cuda(lambda: cq.points_backward(P, Q, ids, g, gq, None), reps)
    res = {"points": int(len(P)), "queries": int(len(Q)), "k": 16, "leaf": 8,
           "backward_query_grad_only_ms": bwd_q_ms,
           "rebin_build_ms": build_ms, "rebin_build_mprims_s": len(P) / build_ms / 1e3,
           "num_nodes": int(bvh.numNodes),
           "knn_ms": knn_ms, "knn_mqueries_s": len(Q) / knn_ms / 1e3,
           "backward_ms": bwd_ms, "fwd_bwd_mqueries_s": len(Q) / (knn_ms + bwd_ms) / 1e3,
           "knn_roofline": roof}
    bvh.free()
    return res


def build_model_bytes(cq, bvh, n):
    """Algorithmic bytes of one spatial-median build, counted from its tree:
    24 B box read + 4 B primID written per prim, 32 B per node written, a 12 B
    centroid written per prim, 20 B per prim and tree level while its node is
    open (12 B centroid read, 4 B node id read and written) and a 24 B leaf
    box read for the bounds.  Returns (bytes, sum over levels of open prims)."""
    import numpy as np
    oc = cq.nodes_to_numpy(bvh)["oc"].astype(np.uint64)
    off = (oc & np.uint64((1 << 48) - 1)).astype(np.int64)
    cnt = (oc >> np.uint64(48)).astype(np.int64)
    frontier, depth, open_prims = np.array([0], np.int64), 0, 0
    while frontier.size:
        leaf = cnt[frontier] > 0
        open_prims += int(cnt[frontier[leaf]].sum()) * depth
        inner = frontier[~leaf]
        frontier = np.concatenate([off[inner], off[inner] + 1])
        depth += 1
    m = int(bvh.numNodes)
    return 64 * n + 32 * m + 20 * open_prims, open_prims


def build_kernel_roofline(cq, boxes, cfg, mem, kern):
    """Per-kernel timing of the spatial-median build (library HIP events on
    the build stream) and the roofline of its dominant kernel, sm::k_pair."""
    import torch
    cq.profile.reset()
    cq.profile.enable(True)
    for _ in range(3):
        cq.gpuBuilder(boxes, cfg, mem=mem).free()
    torch.cuda.synchronize()
    cq.profile.enable(False)
    for name in ["sm_partition", "sm_bounds", "sm_small"]:
        st = cq.profile.query(name)
        if st and st["launches"]:
            kern[name] = {"launches_per_build": st["launches"] / 3,
                          "avg_launch_ms": st["ms"] / st["launches"],
                          "ms_per_build": st["ms"] / 3,
                          "bytes_per_launch": st["bytes"] / st["launches"],
                          "gbs": st["bytes"] / (st["ms"] * 1e-3) / 1e9}
    dom = kern.get("sm_partition")
    if not dom:
        return None
    return {"bound": "hbm", "kernel": "sm_partition",
            "achieved": dom["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": dom["gbs"] / HBM_PEAK_GBS, "traffic": partition_traffic(),
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
            "bytes_model": "32 B per prim of the pass (16-byte item read + written; "
                           "one pass places the prims two tree levels down)",
            "avg_launch_ms": dom["avg_launch_ms"],
            "bytes_per_launch": dom["bytes_per_launch"]}


def bench_radix(cq, boxes, cfg, mem, N, steps, kern):
    """Morton builders: 30-bit keys per the workload, 63-bit reference default."""
    import torch
    radix = {}
    for name, fn in [("radix30", cq.cuda.radix30Builder), ("radix63", cq.cuda.radixBuilder)]:
        fn(boxes, cfg, mem=mem).free()
        torch.cuda.synchronize()
        r0 = time.perf_counter()
        reps = max(3, steps // 2)
        for _ in range(reps):
            fn(boxes, cfg, mem=mem).free()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - r0) / reps * 1e3
        radix[name] = {"ms_per_build": ms, "mprims_s": N / (ms * 1e-3) / 1e6}
    print(f"[bench] radix30 {radix['radix30']['mprims_s']:.1f} Mprims/s, "
          f"radix63 {radix['radix63']['mprims_s']:.1f} Mprims/s", file=sys.stderr, flush=True)
    cq.profile.reset()
    cq.profile.enable(True)
    for _ in range(2):
        cq.cuda.radix30Builder(boxes, cfg, mem=mem).free()
    torch.cuda.synchronize()
    cq.profile.enable(False)
    for name in ["radix_keys", "radix_onesweep", "radix_hierarchy", "radix_emit", "refit"]:
        st = cq.profile.query(name)
        if st and st["launches"]:
            kern["radix30:" + name] = {"launches_per_build": st["launches"] / 2,
                                       "avg_launch_ms": st["ms"] / st["launches"],
                                       "ms_per_build": st["ms"] / 2,
                                       "gbs": st["bytes"] / (st["ms"] * 1e-3) / 1e9}
    return radix


def bench_queries(cq, P, Q, boxes, cfg, mem, N, nq, nk, build_ms):
    """10M closest-point and 1M kNN (k=8) queries on one tree, their
    visited-bytes rooflines, and the whole build against HBM peak."""
    import torch
    bvh = cq.gpuBuilder(boxes, cfg, mem=mem)
    Qf, Qk = Q[:nq], Q[:nk]
    fcp_ms = _time_cuda(lambda: cq.points.findClosest(bvh, P, Qf), 3)
    knn_ms = _time_cuda(lambda: cq.points.findKNN(bvh, P, Qk, 8), 3)
    fcp_kms = _kernel_ms(cq, "fcp_points", lambda: cq.points.findClosest(bvh, P, Qf))
    knn_kms = _kernel_ms(cq, "knn_points", lambda: cq.points.findKNN(bvh, P, Qk, 8))
    query_rooflines = {
        # query 12 B in, id + distance 8 B out; 4 B primID + 12 B point per prim
        "fcp": query_roofline(cq, cq.query_stats(bvh, Qf, data=P, k=0), fcp_kms, 20, 16, "fcp_points"),
        "knn8": query_roofline(cq, cq.query_stats(bvh, Qk, data=P, k=8), knn_kms, 12 + 8 * 8 + 4, 16,
                               "knn_points<8>"),
    }
    num_nodes = bvh.numNodes
    model_bytes, open_prims = build_model_bytes(cq, bvh, N)
    bvh.free()
    torch.cuda.synchronize()
    build_roofline = None
    if build_ms:
        # the whole build against HBM peak, with the per-level byte model
        gbs = model_bytes / (build_ms * 1e-3) / 1e9
        build_roofline = {"bound": "hbm", "bytes": model_bytes, "sum_open_prims": open_prims,
                          "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                          "bytes_model": "64 B/prim (box read, primID, centroid, leaf box) + 32 B/node "
                                         "+ 20 B per prim and level in an open node"}
    return fcp_ms, knn_ms, query_rooflines, num_nodes, build_roofline


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--points", type=int, default=10_000_000)
    ap.add_argument("--fcp-queries", type=int, default=10_000_000)
    ap.add_argument("--knn-queries", type=int, default=1_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=2_000_000)
    ap.add_argument("--cpu-min-s", type=float, default=10.0)
    ap.add_argument("--batch-clouds", type=int, default=4096)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--plain-budget-s", type=float, default=40.0)
    ap.add_argument("--no-extra", action="store_true", help="skip the triangle and mixture legs")
    ap.add_argument("--legs", default="build,radix,queries,triangles,mixture,batch,cpu",
                    help="comma list of legs to run (profiling runs pick a few); value and "
                         "ms_per_step come from the build leg")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 5 device copies of 1 GiB (reference for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "sharding plan is described in DESIGN.md and is not part of this build.",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import cubql_amd as cq
    from cubql_amd import synth

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.calibrate_copy:
        # a plain copy of known size: each dispatch reads and writes 1 GiB
        x = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_dispatch": x.numel() * 4, "dispatches": 5}))
        return
    N = args.points
    P = synth.lcg_uniform_torch(2, N, device=dev)
    boxes = cq.points_to_boxes(P)
    Q = synth.lcg_uniform_torch(3, max(args.fcp_queries, args.knn_queries), device=dev)
    cfg = cq.BuildConfig(makeLeafThreshold=8)
    mem = cq.TorchMemory(dev)
    torch.cuda.synchronize()
    print(f"[bench] data ready: {N:,} points", file=sys.stderr, flush=True)

    legs = set(args.legs.split(","))
    if args.no_extra:
        legs -= {"triangles", "mixture"}
    if args.no_batch:
        legs.discard("batch")
    if args.no_cpu_baseline:
        legs.discard("cpu")
    value = ms_per_step = roofline = None
    kern, radix = {}, {}
    if "build" in legs:
        for _ in range(args.warmup):
            cq.gpuBuilder(boxes, cfg, mem=mem).free()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            bvh = cq.gpuBuilder(boxes, cfg, mem=mem)
            bvh.free()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ms_per_step = (t1 - t0) / args.steps * 1e3
        value = N / (ms_per_step * 1e-3) / 1e6
        print(f"[bench] build: {ms_per_step:.3f} ms/build, {value:.1f} Mprims/s", file=sys.stderr,
              flush=True)
        roofline = build_kernel_roofline(cq, boxes, cfg, mem, kern)
    if "radix" in legs:
        radix = bench_radix(cq, boxes, cfg, mem, N, args.steps, kern)
    tris = mixture = batched = cpu = None
    query_rooflines = None
    fcp_ms = knn_ms = None
    nq, nk = args.fcp_queries, args.knn_queries
    num_nodes = build_roofline = None
    if "queries" in legs:
        (fcp_ms, knn_ms, query_rooflines, num_nodes,
         build_roofline) = bench_queries(cq, P, Q, boxes, cfg, mem, N, nq, nk, ms_per_step)
        print(f"[bench] fcp {nq/fcp_ms/1e3:.1f} Mq/s, knn {nk/knn_ms/1e3:.1f} Mq/s", file=sys.stderr,
              flush=True)
    if "triangles" in legs:
        tris = bench_triangles(cq, dev, reps=5)
        print(f"[bench] triangles: SAH build {tris['sah_build_ms']:.2f} ms, fcp "
              f"{tris['fcp_mqueries_s']:.1f} Mq/s, fwd+bwd {tris['fwd_bwd_mqueries_s']:.1f} Mq/s",
              file=sys.stderr, flush=True)
    if "mixture" in legs:
        mixture = bench_mixture(cq, dev, reps=5)
        print(f"[bench] mixture: rebin build {mixture['rebin_build_ms']:.2f} ms, knn16 "
              f"{mixture['knn_mqueries_s']:.1f} Mq/s, fwd+bwd {mixture['fwd_bwd_mqueries_s']:.1f} Mq/s",
              file=sys.stderr, flush=True)
    if "batch" in legs:
        batched = bench_batched(cq, dev, args.batch_clouds, args.plain_budget_s, not args.no_tune)
        print(f"[bench] batched chamfer: {batched['batched_ms']:.2f} ms, plain {batched['plain_ms']:.1f} ms"
              f" ({batched['speedup_vs_plain']:.1f}x)", file=sys.stderr, flush=True)

    # host->device rate for callers that hold host arrays (never part of value)
    hbox = torch.empty(boxes.shape, dtype=torch.float32, pin_memory=True)
    hbox.copy_(boxes)
    dbox = torch.empty_like(boxes)
    dbox.copy_(hbox, non_blocking=True)
    torch.cuda.synchronize()
    h2d_ms = _time_cuda(lambda: dbox.copy_(hbox, non_blocking=True), 5)
    pcie_h2d_gbs = hbox.numel() * 4 / (h2d_ms * 1e-3) / 1e9
    del hbox, dbox

    if "cpu" in legs:
        cpu = cpu_baseline(min(args.cpu_sample, N), 200_000, min_s=args.cpu_min_s)
        print(f"[bench] cpu baseline {cpu['value']:.2f} Mprims/s", file=sys.stderr, flush=True)

    line = {
        "metric": "build Mprims/s @10M pts; fcp/kNN Mqueries/s; batched fwd+bwd vs plain",
        "value": value,
        "unit": "Mprims/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / PUBLISHED_MPRIMS if value else None,
        "dtype": "f32",
        "data": "synthetic: reference LCG<8> uniform generator, points seed 2, queries seed 3",
        "config": {"workload": f"gpuBuilder (spatial median, leaf<=8) over {N:,} uniform float3 "
                               f"points in [0,1]^3; {nq:,} closest-point and {nk:,} kNN (k=8) "
                               "queries uniform in [0,1]^3",
                   "points": N, "leaf_threshold": 8, "num_nodes": num_nodes},
        "radix_builds": radix,
        "fcp_mqueries_s": nq / fcp_ms / 1e3 if fcp_ms else None,
        "knn_mqueries_s": nk / knn_ms / 1e3 if knn_ms else None,
        "query_rooflines": query_rooflines,
        "pcie_h2d_gbs": pcie_h2d_gbs,
        "triangle_mesh": tris,
        "gaussian_mixture": mixture,
        "batched_chamfer": batched,
        "kernels": kern,
        "roofline": roofline,
        "build_roofline": build_roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
