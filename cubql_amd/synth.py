"""Synthetic workload generators (host side, numpy).

``lcg_uniform`` reproduces the reference's UniformPointGenerator: one
``LCG<8>(seed, index)`` per point, D draws of 24-bit floats, coordinates
computed in double and rounded to float32 (math/random.h:24-68,
samples/common/Generator.cpp:202-218, samples/common/IO.h:61-66).
"""
from __future__ import annotations

import numpy as np

_M32 = np.uint64(0xFFFFFFFF)


def _u32(x):
    return np.asarray(x, dtype=np.uint64) & _M32


def lcg_init(seed: int, index: np.ndarray) -> np.ndarray:
    """LCG<8>::init(seed, index): 8 rounds of the TEA-style mixing (random.h:31-44)."""
    v0 = np.full(index.shape, seed & 0xFFFFFFFF, dtype=np.uint64)
    v1 = _u32(index.astype(np.uint64))
    s0 = np.uint64(0)
    for _ in range(8):
        s0 = (s0 + np.uint64(0x9E3779B9)) & _M32
        v0 = (v0 + ((((v1 << np.uint64(4)) & _M32) + np.uint64(0xA341316C)) & _M32
                    ^ ((v1 + s0) & _M32)
                    ^ (((v1 >> np.uint64(5)) + np.uint64(0xC8013EA4)) & _M32))) & _M32
        v1 = (v1 + ((((v0 << np.uint64(4)) & _M32) + np.uint64(0xAD90777D)) & _M32
                    ^ ((v0 + s0) & _M32)
                    ^ (((v0 >> np.uint64(5)) + np.uint64(0x7E95761E)) & _M32))) & _M32
    return v0


def lcg_uniform(seed: int, n: int, lo: float = 0.0, hi: float = 1.0, dims: int = 3,
                chunk: int = 1 << 22) -> np.ndarray:
    """[n, dims] float32 points, identical to UniformPointGenerator<dims>."""
    out = np.empty((n, dims), dtype=np.float32)
    lo_d, span = np.float64(lo), np.float64(hi) - np.float64(lo)
    for b in range(0, n, chunk):
        e = min(n, b + chunk)
        state = lcg_init(seed, np.arange(b, e, dtype=np.uint64))
        for d in range(dims):
            state = (state * np.uint64(1664525) + np.uint64(1013904223)) & _M32
            r = (state & np.uint64(0x00FFFFFF)).astype(np.float32) / np.float32(16777216.0)
            out[b:e, d] = (lo_d + r.astype(np.float64) * span).astype(np.float32)
    return out


def points_to_boxes(p: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(p, dtype=np.float32).reshape(-1, 3)
    return np.ascontiguousarray(np.concatenate([p, p], axis=1))


def lcg_uniform_torch(seed: int, n: int, lo: float = 0.0, hi: float = 1.0, dims: int = 3,
                      device="cuda"):
    """Same values as lcg_uniform, computed on the device with int64 tensor ops."""
    import torch
    M = 0xFFFFFFFF
    idx = torch.arange(n, dtype=torch.int64, device=device)
    v0 = torch.full((n,), seed & M, dtype=torch.int64, device=device)
    v1 = idx & M
    s0 = 0
    for _ in range(8):
        s0 = (s0 + 0x9E3779B9) & M
        v0 = (v0 + ((((v1 << 4) & M) + 0xA341316C) & M ^ ((v1 + s0) & M)
                    ^ (((v1 >> 5) + 0xC8013EA4) & M))) & M
        v1 = (v1 + ((((v0 << 4) & M) + 0xAD90777D) & M ^ ((v0 + s0) & M)
                    ^ (((v0 >> 5) + 0x7E95761E) & M))) & M
    state = v0
    out = torch.empty((n, dims), dtype=torch.float32, device=device)
    span = float(hi) - float(lo)
    for d in range(dims):
        state = (state * 1664525 + 1013904223) & M
        r = (state & 0x00FFFFFF).to(torch.float32) / 16777216.0
        out[:, d] = (float(lo) + r.to(torch.float64) * span).to(torch.float32)
    return out
