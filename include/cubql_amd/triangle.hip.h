// triangle.hip.h -- closest point on one triangle (device code).
//
// The per-triangle test of cuBQL::triangles::closestPoint
// (queries/triangles/findClosest.h:56-122): N = (b-a)x(c-a); the query is
// outside if N == 0 or it lies beyond any edge plane, and then the closest
// of the three clamped edge projections wins (first minimal edge on equal
// distances); otherwise the plane distance along normalize(N) =
// N*(1/sqrt(N.N)).  Float32 with contraction off, so distances round the
// same way in every caller (batched kernels, the adjoint, the device API).
// Optionally returns the barycentric weights of the closest point (face:
// projection barycentrics, edge: (1-t, t), vertex: one-hot).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace cubql_dev {

struct F3 { float x, y, z; };
__device__ __forceinline__ F3 ld3(const float *p, size_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 mul(float s, F3 a) { return {s * a.x, s * a.y, s * a.z}; }
// dot (vec.h:458-466) and cross (vec.h:232-237), no contraction
__device__ __forceinline__ float dot(F3 a, F3 b) {
#pragma clang fp contract(off)
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
#pragma clang fp contract(off)
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ float clamp01(float t) { return fminf(fmaxf(t, 0.f), 1.f); }

__device__ __forceinline__ void tri_verts(const int32_t *tris, const float *verts, uint32_t t, F3 &a,
                                          F3 &b, F3 &c) {
  if (tris) {
    a = ld3(verts, (size_t)tris[3 * (size_t)t + 0]);
    b = ld3(verts, (size_t)tris[3 * (size_t)t + 1]);
    c = ld3(verts, (size_t)tris[3 * (size_t)t + 2]);
  } else {
    a = ld3(verts, 3 * (size_t)t);
    b = ld3(verts, 3 * (size_t)t + 1);
    c = ld3(verts, 3 * (size_t)t + 2);
  }
}

__device__ __forceinline__ float edge_closest(F3 a, F3 b, F3 p, float &t, F3 &cp) {
  F3 ab = sub(b, a);
  t = clamp01(dot(sub(p, a), ab) / dot(ab, ab));
  cp = add(a, mul(t, ab));
  F3 d = sub(cp, p);
  return dot(d, d);
}

// closestPoint(Triangle) (findClosest.h:93-122); returns the squared
// distance, the closest point P and (when WANT_W) its barycentric weights.
template <bool WANT_W>
__device__ __forceinline__ float tri_closest(F3 a, F3 b, F3 c, F3 q, F3 &P, float (&w)[3]) {
  const F3 N = cross(sub(b, a), sub(c, a));
  const bool outside = (N.x == 0.f && N.y == 0.f && N.z == 0.f) ||
                       (dot(sub(q, a), cross(sub(b, a), N)) >= 0.f) ||
                       (dot(sub(q, b), cross(sub(c, b), N)) >= 0.f) ||
                       (dot(sub(q, c), cross(sub(a, c), N)) >= 0.f);
  if (outside) {
    float t0, t1, t2;
    F3 c0, c1, c2;
    const float d0 = edge_closest(a, b, q, t0, c0);
    const float d1 = edge_closest(b, c, q, t1, c1);
    const float d2 = edge_closest(c, a, q, t2, c2);
    // sequential strictly-closer acceptance: first minimal edge wins
    float best = d0;
    P = c0;
    int e = 0;
    if (d1 < best) { best = d1; P = c1; e = 1; }
    if (d2 < best) { best = d2; P = c2; e = 2; }
    if (WANT_W) {
      if (e == 0) { w[0] = 1.f - t0; w[1] = t0; w[2] = 0.f; }
      else if (e == 1) { w[0] = 0.f; w[1] = 1.f - t1; w[2] = t1; }
      else { w[0] = t2; w[1] = 0.f; w[2] = 1.f - t2; }
    }
    return best;
  }
  const float nn = dot(N, N);
  const F3 Nn = mul(1.f / sqrtf(nn), N);
  const float sd = dot(sub(q, a), Nn);
  P = sub(q, mul(sd, Nn));
  if (WANT_W) {
    const float wa = dot(cross(sub(c, b), sub(P, b)), N) / nn;
    const float wb = dot(cross(sub(a, c), sub(P, c)), N) / nn;
    w[0] = wa;
    w[1] = wb;
    w[2] = 1.f - wa - wb;
  }
  return sd * sd;
}

}  // namespace cubql_dev
