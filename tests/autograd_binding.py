"""torch.autograd.Function bindings of the distance adjoints (test harness only).

The product exposes the adjoints as C-ABI entry points
(cubql_points_distance_backward_f3, cubql_fcp_triangles_backward_f3); these
thin Functions put them behind torch's autograd so that
torch.autograd.gradcheck can compare them with finite differences:

* forward builds a BVH over the float32 copy of the data, runs the HIP
  query for the neighbour / triangle ids (exact integer work, never
  differentiated) and returns the distances of those ids;
* backward calls the HIP adjoint kernel with the saved ids and returns its
  float32 gradients in the inputs' dtype.

With float64 inputs the forward evaluates the distances of the HIP ids in
float64 (points: |q - p|; triangles: closest point on the triangle by the
textbook region method, Ericson, Real-Time Collision Detection 5.1.5), so
gradcheck's float64 finite differences measure the float32 adjoint kernel
against the exact derivative.  With float32 inputs the distances are the
HIP query's own.
"""
from __future__ import annotations

import torch

import cubql_amd as cq


def _f32(t):
    return t.detach().to(torch.float32).contiguous()


class PointDistance(torch.autograd.Function):
    """d[i, j] = |queries[i] - points[ids[i, j]]| over the k nearest points
    (k = 1: closest point).  Returns (d [nq, k], ids [nq, k])."""

    @staticmethod
    def forward(ctx, queries, points, k: int, leaf: int = 8):
        q32, p32 = _f32(queries), _f32(points)
        bvh = cq.gpuBuilder(cq.points_to_boxes(p32), cq.BuildConfig(makeLeafThreshold=leaf))
        if k == 1:
            ids, d2 = cq.points.findClosest(bvh, p32, q32)
            ids, d2 = ids.view(-1, 1), d2.view(-1, 1)
        else:
            ids, d2, _ = cq.points.findKNN(bvh, p32, q32, k)
        torch.cuda.synchronize()
        bvh.free()
        valid = ids >= 0
        if queries.dtype == torch.float64:
            diff = queries.detach()[:, None, :] - points.detach()[ids.clamp(min=0).long()]
            d = torch.where(valid, diff.norm(dim=-1), torch.zeros((), dtype=torch.float64,
                                                                  device=queries.device))
        else:
            d = torch.where(valid, torch.sqrt(d2), torch.zeros_like(d2))
        ctx.save_for_backward(queries, points, ids)
        ctx.mark_non_differentiable(ids)
        return d, ids

    @staticmethod
    def backward(ctx, grad_d, _grad_ids):
        queries, points, ids = ctx.saved_tensors
        q32, p32 = _f32(queries), _f32(points)
        gq = torch.empty_like(q32)
        gp = torch.zeros_like(p32)
        k = ids.shape[1]
        cq.points_backward(p32, q32, ids if k > 1 else ids.view(-1), _f32(grad_d).view(ids.shape),
                           gq, gp)
        return gq.to(queries.dtype), gp.to(points.dtype), None, None


def closest_point_on_triangle_f64(p, a, b, c):
    """Closest point of each triangle (a, b, c) to p, by Voronoi regions
    (vertex, edge, face), in the inputs' dtype."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = (ab * ap).sum(-1), (ac * ap).sum(-1)
    res = a.clone()
    done = (d1 <= 0) & (d2 <= 0)
    bp = p - b
    d3, d4 = (ab * bp).sum(-1), (ac * bp).sum(-1)
    m = ~done & (d3 >= 0) & (d4 <= d3)
    res[m] = b[m]
    done |= m
    vc = d1 * d4 - d3 * d2
    m = ~done & (vc <= 0) & (d1 >= 0) & (d3 <= 0)
    v = d1 / (d1 - d3)
    res[m] = (a + v[:, None] * ab)[m]
    done |= m
    cp = p - c
    d5, d6 = (ab * cp).sum(-1), (ac * cp).sum(-1)
    m = ~done & (d6 >= 0) & (d5 <= d6)
    res[m] = c[m]
    done |= m
    vb = d5 * d2 - d1 * d6
    m = ~done & (vb <= 0) & (d2 >= 0) & (d6 <= 0)
    w = d2 / (d2 - d6)
    res[m] = (a + w[:, None] * ac)[m]
    done |= m
    va = d3 * d6 - d5 * d4
    m = ~done & (va <= 0) & ((d4 - d3) >= 0) & ((d5 - d6) >= 0)
    w = (d4 - d3) / ((d4 - d3) + (d5 - d6))
    res[m] = (b + w[:, None] * (c - b))[m]
    done |= m
    denom = 1.0 / (va + vb + vc)
    v, w = vb * denom, vc * denom
    res[~done] = (a + ab * v[:, None] + ac * w[:, None])[~done]
    return res


class TriangleDistance(torch.autograd.Function):
    """d[i] = distance from queries[i] to the closest triangle of the
    indexed mesh (verts, tris).  Returns (d [nq], triangle ids [nq])."""

    @staticmethod
    def forward(ctx, queries, verts, tris, leaf: int = 4):
        q32, v32 = _f32(queries), _f32(verts)
        tris = tris.to(torch.int32).contiguous()
        bvh = cq.gpuBuilder(cq.triangles.boxes(tris, v32), cq.BuildConfig(makeLeafThreshold=leaf))
        ids, d2, _ = cq.triangles.findClosest(bvh, tris, v32, q32, want_points=False)
        torch.cuda.synchronize()
        bvh.free()
        valid = ids >= 0
        if queries.dtype == torch.float64:
            t = tris[ids.clamp(min=0).long()].long()
            V = verts.detach()
            P = closest_point_on_triangle_f64(queries.detach(), V[t[:, 0]], V[t[:, 1]], V[t[:, 2]])
            d = torch.where(valid, (queries.detach() - P).norm(dim=-1),
                            torch.zeros((), dtype=torch.float64, device=queries.device))
        else:
            d = torch.where(valid, torch.sqrt(d2), torch.zeros_like(d2))
        ctx.save_for_backward(queries, verts, tris, ids)
        ctx.mark_non_differentiable(ids)
        return d, ids

    @staticmethod
    def backward(ctx, grad_d, _grad_ids):
        queries, verts, tris, ids = ctx.saved_tensors
        q32, v32 = _f32(queries), _f32(verts)
        gq = torch.empty_like(q32)
        gv = torch.zeros_like(v32)
        cq.triangles.backward(tris, v32, q32, ids, _f32(grad_d), gq, gv)
        return gq.to(queries.dtype), gv.to(verts.dtype), None, None
