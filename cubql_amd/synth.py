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
                chunk: int = 1 << 22, dtype=np.float32) -> np.ndarray:
    """[n, dims] points, identical to UniformPointGenerator<dims>: the
    generator works in double (uniformPointGenerator, Generator.cpp:202-218);
    float32 output is its conversion (samples/common/IO.h:60-66), float64
    output the values a generatePoints .bin file holds."""
    out = np.empty((n, dims), dtype=dtype)
    lo_d, span = np.float64(lo), np.float64(hi) - np.float64(lo)
    for b in range(0, n, chunk):
        e = min(n, b + chunk)
        state = lcg_init(seed, np.arange(b, e, dtype=np.uint64))
        for d in range(dims):
            state = (state * np.uint64(1664525) + np.uint64(1013904223)) & _M32
            r = (state & np.uint64(0x00FFFFFF)).astype(np.float32) / np.float32(16777216.0)
            out[b:e, d] = (lo_d + r.astype(np.float64) * span).astype(dtype)
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


def box_muller(u1: np.ndarray, u2: np.ndarray) -> np.ndarray:
    """Standard normal deviates from two uniform [0,1) streams (float64)."""
    u1 = 1.0 - u1.astype(np.float64)  # (0,1]: log is finite
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2.astype(np.float64))


def lcg_draws(seed: int, n: int, k: int) -> np.ndarray:
    """k successive LCG<8>(seed, i) draws in [0,1) per element i: [n, k]."""
    return lcg_uniform(seed, n, 0.0, 1.0, dims=k)


def icosphere(subdiv: int):
    """Unit icosphere: 20*4^subdiv triangles, 10*4^subdiv+2 vertices (float64)."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    V = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
                  (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)],
                 dtype=np.float64)
    F = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9),
                  (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2),
                  (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
                  (8, 6, 7), (9, 8, 1)], dtype=np.int64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    for _ in range(subdiv):
        a, b, c = F[:, 0], F[:, 1], F[:, 2]
        e = np.concatenate([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)])
        e.sort(axis=1)
        key = e[:, 0] * (len(V) + 1) + e[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        ea, eb = uk // (len(V) + 1), uk % (len(V) + 1)
        mid = V[ea] + V[eb]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(V)
        V = np.concatenate([V, mid])
        m = len(F)
        ab, bc, ca = base + inv[:m], base + inv[m:2 * m], base + inv[2 * m:]
        F = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
                            np.stack([c, ca, bc], 1), np.stack([ab, bc, ca], 1)])
    return V, F


def noisy_icosphere(subdiv: int = 8, seed: int = 4, sigma: float = 0.01):
    """Icosphere of radius 1 with vertices displaced radially by N(0, sigma^2)
    (Box-Muller over LCG<8>(seed, vertex) draws).  Returns (verts f32 [V,3],
    tris i32 [F,3])."""
    V, F = icosphere(subdiv)
    u = lcg_draws(seed, len(V), 2)
    r = 1.0 + sigma * box_muller(u[:, 0], u[:, 1])
    return (V * r[:, None]).astype(np.float32), F.astype(np.int32)


def mesh_surface_queries(verts: np.ndarray, tris: np.ndarray, n: int, seed: int = 5,
                         sigma: float = 0.05) -> np.ndarray:
    """n queries: a triangle chosen with probability proportional to its area,
    a uniform point on it (Triangle::sample folding, Triangle.h:41-45), plus
    an N(0, sigma^2) offset per axis.  All draws from LCG<8>(seed, i)."""
    V = verts.astype(np.float64)
    a, b, c = V[tris[:, 0]], V[tris[:, 1]], V[tris[:, 2]]
    area = np.linalg.norm(np.cross(b - a, c - a), axis=1)
    cdf = np.cumsum(area)
    cdf /= cdf[-1]
    out = np.empty((n, 3), dtype=np.float32)
    chunk = 1 << 21
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        u = _lcg_range(seed, s, e, 9)
        t = np.minimum(np.searchsorted(cdf, u[:, 0], side="right"), len(tris) - 1)
        uu, vv = u[:, 1].copy(), u[:, 2].copy()
        fold = uu + vv >= 1.0
        uu[fold], vv[fold] = 1.0 - uu[fold], 1.0 - vv[fold]
        p = (1.0 - uu - vv)[:, None] * a[t] + uu[:, None] * b[t] + vv[:, None] * c[t]
        g = np.stack([box_muller(u[:, 3 + 2 * j], u[:, 4 + 2 * j]) for j in range(3)], 1)
        out[s:e] = (p + sigma * g).astype(np.float32)
    return out


def _lcg_range(seed: int, b: int, e: int, dims: int) -> np.ndarray:
    """LCG<8>(seed, i) draws for i in [b, e)."""
    state = lcg_init(seed, np.arange(b, e, dtype=np.uint64))
    out = np.empty((e - b, dims), dtype=np.float64)
    for d in range(dims):
        state = (state * np.uint64(1664525) + np.uint64(1013904223)) & _M32
        out[:, d] = ((state & np.uint64(0x00FFFFFF)).astype(np.float32) / np.float32(16777216.0))
    return out


def chamfer_cloud_sizes(num_clouds: int, seed: int, lo: int = 64, hi: int = 65536):
    """Cloud sizes log-uniform in [lo, hi]: one LCG<8>(seed, cloud) draw each."""
    u = lcg_uniform(seed, num_clouds, 0.0, 1.0, dims=1)[:, 0].astype(np.float64)
    return np.floor(np.exp(np.log(lo) + u * (np.log(hi + 1) - np.log(lo)))).astype(np.int64).clip(lo, hi)


def chamfer_clouds_torch(sizes: np.ndarray, seed: int, scale_seed: int, device="cuda"):
    """Concatenated float32 clouds: point i of the batch is N(0, I) from four
    LCG<8>(seed, i) draws (Box-Muller), scaled by its cloud's factor
    U[0.1, 10] (LCG<8>(scale_seed, cloud)).  Returns (points [n,3], offsets
    [numClouds+1] int32), both on `device`."""
    import torch
    n = int(sizes.sum())
    off = np.zeros(len(sizes) + 1, dtype=np.int64)
    off[1:] = np.cumsum(sizes)
    s = 0.1 + lcg_uniform(scale_seed, len(sizes), 0.0, 1.0, dims=1)[:, 0].astype(np.float64) * 9.9
    u = lcg_uniform_torch(seed, n, 0.0, 1.0, dims=4, device=device).to(torch.float64)
    r1 = torch.sqrt(-2.0 * torch.log(1.0 - u[:, 0]))
    r2 = torch.sqrt(-2.0 * torch.log(1.0 - u[:, 2]))
    th1, th2 = 2.0 * np.pi * u[:, 1], 2.0 * np.pi * u[:, 3]
    g = torch.stack([r1 * torch.cos(th1), r1 * torch.sin(th1), r2 * torch.cos(th2)], 1)
    cloud = torch.repeat_interleave(torch.arange(len(sizes), device=device),
                                    torch.from_numpy(sizes).to(device))
    pts = (g * torch.from_numpy(s).to(device)[cloud][:, None]).to(torch.float32).contiguous()
    return pts, torch.from_numpy(off.astype(np.int32)).to(device)


def gaussian_mixture(n: int, seed: int, mixture_seed: int = 40, components: int = 64,
                     outlier_frac: float = 0.01, outlier_radius: float = 1e6) -> np.ndarray:
    """[n,3] float32 points from a fixed Gaussian mixture with outliers.

    The mixture (from `mixture_seed`): `components` centres uniform in
    [-1e3, 1e3]^3 and per-component sigma log-uniform in [1e-4, 1e-1].
    Point i uses eight LCG<8>(seed, i) draws: u0 picks the component, u1 < 
    outlier_frac makes it an outlier at distance outlier_radius from the
    origin in a direction from three normal deviates (u2..u5), otherwise it
    is centre + sigma * N(0, I) (u2..u7, Box-Muller).  Points and queries of
    one workload share `mixture_seed` and differ in `seed`.
    """
    m = lcg_draws(mixture_seed, components, 4).astype(np.float64)
    centres = -1e3 + m[:, :3] * 2e3
    sigma = np.exp(np.log(1e-4) + m[:, 3] * (np.log(1e-1) - np.log(1e-4)))
    u = lcg_draws(seed, n, 8).astype(np.float64)
    comp = np.minimum((u[:, 0] * components).astype(np.int64), components - 1)
    g = np.stack([box_muller(u[:, 2], u[:, 3]), box_muller(u[:, 4], u[:, 5]),
                  box_muller(u[:, 6], u[:, 7])], axis=1)
    pts = centres[comp] + sigma[comp, None] * g
    out = u[:, 1] < outlier_frac
    if out.any():
        d = g[out]
        norm = np.linalg.norm(d, axis=1, keepdims=True)
        d = np.where(norm > 0, d / np.maximum(norm, 1e-300), np.array([1.0, 0.0, 0.0]))
        pts[out] = outlier_radius * d
    return pts.astype(np.float32)
