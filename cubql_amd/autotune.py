"""Autotuner for the batched path: leaf size, query block size, stack depth.

Every candidate computes the same nearest neighbours and distances (the
lowest-index rule makes results independent of the tree shape), so the
tuner only measures time: it replays the captured batch for each candidate
and keeps the fastest.  Results are cached per (device name, batch
signature) in memory and, if `cache_path` is given, in a JSON file.
"""
from __future__ import annotations

import json
import os
import time
from typing import Iterable, Optional

import torch

from . import BatchParams, ChamferBatch

DEFAULT_CANDIDATES = [(T, b, s) for T in (2, 4, 8, 16) for b in (64, 128, 256) for s in (32, 64)]
_cache = {}


def _signature(src, tgt, num_clouds):
    return f"{torch.cuda.get_device_name(src.device)}|c={num_clouds}|s={src.shape[0]}|t={tgt.shape[0]}"


def tune_batch(src: torch.Tensor, src_off: torch.Tensor, tgt: torch.Tensor, tgt_off: torch.Tensor,
               candidates: Optional[Iterable] = None, reps: int = 3,
               cache_path: Optional[str] = None, max_cloud_size: int = 0):
    """Returns (best BatchParams, {candidate: ms}) for this batch."""
    num_clouds = src_off.numel() - 1
    key = _signature(src, tgt, num_clouds)
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            _cache.update(json.load(f))
    if key in _cache:
        T, b, s = _cache[key]["best"]
        return BatchParams(T, b, s, max_cloud_size), _cache[key]["times"]
    times = {}
    for cand in (candidates or DEFAULT_CANDIDATES):
        p = BatchParams(*cand, max_cloud_size)
        batch = ChamferBatch(src, src_off, tgt, tgt_off, p).capture()
        batch.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            batch.replay()
        torch.cuda.synchronize()
        times[str(tuple(cand))] = (time.perf_counter() - t0) / reps * 1e3
        del batch
    best = min(times, key=times.get)
    T, b, s = (int(x) for x in best.strip("()").split(","))
    _cache[key] = {"best": [T, b, s], "times": times}
    if cache_path:
        with open(cache_path, "w") as f:
            json.dump(_cache, f, indent=1)
    return BatchParams(T, b, s, max_cloud_size), times
