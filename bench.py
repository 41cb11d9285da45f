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


def cpu_baseline(n_sample: int, nq_sample: int, budget_s: float):
    """Reference CPU path (cpu::spatialMedian + host findClosest), 1 core."""
    import numpy as np

    import oracle
    from cubql_amd import synth

    P = synth.lcg_uniform(2, n_sample)
    Q = synth.lcg_uniform(3, nq_sample)
    boxes = oracle.points_to_boxes(P)
    kind = "reference" if oracle.ref_available() else "port"
    if kind == "reference":
        R = oracle.ref()
        build_once = lambda: R.ref_time_cpu_build(oracle._p(boxes), n_sample, 8)  # noqa: E731
    else:
        def build_once():
            t = time.perf_counter()
            oracle.sm_build(boxes, 8)
            return time.perf_counter() - t
    build_once()  # warm-up (buildBench.h:115-117)
    times = []
    t_start = time.perf_counter()
    while len(times) < 3 or (time.perf_counter() - t_start < budget_s and len(times) < 10):
        times.append(build_once())
    tree = oracle.ref_sm_build(boxes, 8) if kind == "reference" else oracle.sm_build(boxes, 8)
    t = time.perf_counter()
    if kind == "reference":
        oracle.ref_fcp(tree, P, Q)
    else:
        oracle.fcp(tree, P, Q, lowest_id=False)
    tq = time.perf_counter() - t
    cores, model = cpu_info()
    mean = float(np.mean(times))
    return {
        "value": n_sample / mean / 1e6,
        "unit": "Mprims/s",
        "cores": 1,
        "kind": kind,
        "sample": (f"cpu::spatialMedian (serial, leaf<=8) over the first {n_sample:,} points of the "
                   f"same seed-2 stream, {len(times)} timed builds after 1 warm-up "
                   f"(mean {mean*1e3:.1f} ms, median {float(np.median(times))*1e3:.1f} ms); "
                   f"host findClosest over {nq_sample:,} seed-3 queries on that tree"),
        "fcp_mqueries_s": nq_sample / tq / 1e6,
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
        cands = [(T, b, st) for T in (4, 8, 16) for b in (128, 256) for st in (32, 64)]
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
           "sah_kernels": kern}
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
    ids = out["r"][0]
    g = torch.full((len(Q), 16), 1.0 / (len(Q) * 16), device=dev)
    gq = torch.empty_like(Q)
    gp = torch.zeros_like(P)
    bwd_ms = _time_cuda(lambda: cq.points_backward(P, Q, ids, g, gq, gp), reps)
    res = {"points": int(len(P)), "queries": int(len(Q)), "k": 16, "leaf": 8,
           "rebin_build_ms": build_ms, "rebin_build_mprims_s": len(P) / build_ms / 1e3,
           "num_nodes": int(bvh.numNodes),
           "knn_ms": knn_ms, "knn_mqueries_s": len(Q) / knn_ms / 1e3,
           "backward_ms": bwd_ms, "fwd_bwd_mqueries_s": len(Q) / (knn_ms + bwd_ms) / 1e3}
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
    ap.add_argument("--batch-clouds", type=int, default=4096)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--plain-budget-s", type=float, default=40.0)
    ap.add_argument("--no-extra", action="store_true", help="skip the triangle and mixture legs")
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

    # ---- per-kernel roofline of the dominant build kernel (events in the library)
    cq.profile.reset()
    cq.profile.enable(True)
    for _ in range(3):
        cq.gpuBuilder(boxes, cfg, mem=mem).free()
    torch.cuda.synchronize()
    cq.profile.enable(False)
    part = cq.profile.query("sm_partition")
    bnds = cq.profile.query("sm_bounds")
    small = cq.profile.query("sm_small")
    kern = {}
    for name, st in [("sm_partition", part), ("sm_bounds", bnds), ("sm_small", small)]:
        if st and st["launches"]:
            kern[name] = {"launches_per_build": st["launches"] / 3,
                          "avg_launch_ms": st["ms"] / st["launches"],
                          "ms_per_build": st["ms"] / 3,
                          "bytes_per_launch": st["bytes"] / st["launches"],
                          "gbs": st["bytes"] / (st["ms"] * 1e-3) / 1e9}
    dom = kern.get("sm_partition")
    roofline = None
    if dom:
        roofline = {"bound": "hbm", "kernel": "sm_partition",
                    "achieved": dom["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": dom["gbs"] / HBM_PEAK_GBS, "traffic": partition_traffic(),
                    "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
                    "bytes_model": "32 B per prim of the pass (16-byte item read + written; "
                                   "one pass places the prims two tree levels down)",
                    "avg_launch_ms": dom["avg_launch_ms"],
                    "bytes_per_launch": dom["bytes_per_launch"]}

    # ---- Morton builders (30-bit keys per the workload, 63-bit reference default)
    radix = {}
    for name, fn in [("radix30", cq.cuda.radix30Builder), ("radix63", cq.cuda.radixBuilder)]:
        fn(boxes, cfg, mem=mem).free()
        torch.cuda.synchronize()
        r0 = time.perf_counter()
        reps = max(3, args.steps // 2)
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
    for name in ["radix_keys", "radix_onesweep", "radix_hierarchy", "refit"]:
        st = cq.profile.query(name)
        if st and st["launches"]:
            kern["radix30:" + name] = {"launches_per_build": st["launches"] / 2,
                                       "avg_launch_ms": st["ms"] / st["launches"],
                                       "ms_per_build": st["ms"] / 2,
                                       "gbs": st["bytes"] / (st["ms"] * 1e-3) / 1e9}

    # ---- queries on one tree
    bvh = cq.gpuBuilder(boxes, cfg, mem=mem)
    nq = args.fcp_queries
    Qf = Q[:nq]
    cq.points.findClosest(bvh, P, Qf)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 3
    e0.record()
    for _ in range(reps):
        cq.points.findClosest(bvh, P, Qf)
    e1.record()
    torch.cuda.synchronize()
    fcp_ms = e0.elapsed_time(e1) / reps
    nk = args.knn_queries
    Qk = Q[:nk]
    cq.points.findKNN(bvh, P, Qk, 8)
    e0.record()
    for _ in range(reps):
        cq.points.findKNN(bvh, P, Qk, 8)
    e1.record()
    torch.cuda.synchronize()
    knn_ms = e0.elapsed_time(e1) / reps
    num_nodes = bvh.numNodes
    model_bytes, open_prims = build_model_bytes(cq, bvh, N)
    bvh.free()
    # the whole build against HBM peak, with the per-level byte model
    build_roofline = {"bound": "hbm", "bytes": model_bytes, "sum_open_prims": open_prims,
                      "achieved": model_bytes / (ms_per_step * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": model_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
                      "bytes_model": "64 B/prim (box read, primID, centroid, leaf box) + 32 B/node "
                                     "+ 20 B per prim and level in an open node"}
    print(f"[bench] fcp {nq/fcp_ms/1e3:.1f} Mq/s, knn {nk/knn_ms/1e3:.1f} Mq/s", file=sys.stderr,
          flush=True)

    tris = mixture = None
    if not args.no_extra:
        tris = bench_triangles(cq, dev, reps=5)
        print(f"[bench] triangles: SAH build {tris['sah_build_ms']:.2f} ms, fcp "
              f"{tris['fcp_mqueries_s']:.1f} Mq/s, fwd+bwd {tris['fwd_bwd_mqueries_s']:.1f} Mq/s",
              file=sys.stderr, flush=True)
        mixture = bench_mixture(cq, dev, reps=5)
        print(f"[bench] mixture: rebin build {mixture['rebin_build_ms']:.2f} ms, knn16 "
              f"{mixture['knn_mqueries_s']:.1f} Mq/s, fwd+bwd {mixture['fwd_bwd_mqueries_s']:.1f} Mq/s",
              file=sys.stderr, flush=True)

    batched = None
    if not args.no_batch:
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

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(min(args.cpu_sample, N), 200_000, budget_s=12.0)
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
        "vs_baseline": value / PUBLISHED_MPRIMS,
        "dtype": "f32",
        "data": "synthetic: reference LCG<8> uniform generator, points seed 2, queries seed 3",
        "config": {"workload": f"gpuBuilder (spatial median, leaf<=8) over {N:,} uniform float3 "
                               f"points in [0,1]^3; {nq:,} closest-point and {nk:,} kNN (k=8) "
                               "queries uniform in [0,1]^3",
                   "points": N, "leaf_threshold": 8, "num_nodes": num_nodes},
        "radix_builds": radix,
        "fcp_mqueries_s": nq / fcp_ms / 1e3,
        "knn_mqueries_s": nk / knn_ms / 1e3,
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
