// cuBQL/queries/triangles/findClosest.h -- device closest-point queries on
// triangle meshes (queries/triangles/findClosest.h:27-152).
//
//   struct FCPResult { int primID; float sqrDist; vec3f P; clear(maxSqr) }
//   bool closestPoint(FCPResult &, Triangle, vec3f queryPoint)
//   void findClosest(FCPResult &, bvh3f, const Triangle *, vec3f queryPoint)
//   void fcp(FCPResult &, vec3f queryPoint, bvh3f, const vec3i *triangles,
//            const vec3f *vertices)
//
// The per-triangle test is the one the library's batched kernels use
// (cubql_amd/triangle.hip.h), so distances and closest points are
// bit-identical to cubql_fcp_triangles_f3.  Among equal distances the
// lowest triangle id wins (the reference keeps the first one found).
// Clear the result with result.clear(maxSqrDist) before a query; only
// triangles within that squared distance are accepted.
#pragma once
#include "cuBQL/queries/triangles/Triangle.h"
#include "cuBQL/traversal/shrinkingRadiusQuery.h"
#include "cubql_amd/triangle.hip.h"

namespace cuBQL {
namespace triangles {

struct FCPResult {
  __device__ inline void clear(float maxDistSqr) {
    primID = -1;
    sqrDist = maxDistSqr;
  }
  int primID;
  float sqrDist;
  vec3f P;
};

namespace detail {
__device__ inline cubql_dev::F3 f3(vec3f v) { return {v.x, v.y, v.z}; }
// squared distance and closest point of one triangle
__device__ inline float test(const Triangle &t, vec3f q, vec3f &P) {
  cubql_dev::F3 cp;
  float w[3];
  const float d = cubql_dev::tri_closest<false>(f3(t.a), f3(t.b), f3(t.c), f3(q), cp, w);
  P = vec3f(cp.x, cp.y, cp.z);
  return d;
}
}  // namespace detail

/*! closestPoint (findClosest.h:93-122): updates result and returns true if
    the triangle is strictly closer than result.sqrDist. */
__device__ inline bool closestPoint(FCPResult &result, Triangle triangle, vec3f qp) {
  vec3f P;
  const float d = detail::test(triangle, qp, P);
  if (d >= result.sqrDist) return false;
  result.sqrDist = d;
  result.P = P;
  return true;
}

__device__ inline void findClosest(FCPResult &result, const bvh3f bvh, const Triangle *triangles,
                                   const vec3f queryPoint) {
  auto perPrim = [&](uint32_t primID) -> float {
    vec3f P;
    const float d = detail::test(triangles[primID], queryPoint, P);
    if (d < result.sqrDist || (d == result.sqrDist && result.primID >= 0 && (int)primID < result.primID)) {
      result.sqrDist = d;
      result.primID = (int)primID;
      result.P = P;
    }
    return result.sqrDist;
  };
  shrinkingRadiusQuery::forEachPrim(perPrim, bvh, queryPoint, result.sqrDist);
}

/*! fcp (findClosest.h:43-48): the indexed-mesh form. */
__device__ inline void fcp(FCPResult &result, const vec3f queryPoint, const bvh3f bvh,
                           const vec3i *const __restrict__ triangles, const vec3f *const __restrict__ vertices) {
  auto perPrim = [&](uint32_t primID) -> float {
    const vec3i t = triangles[primID];
    const Triangle tri{vertices[t.x], vertices[t.y], vertices[t.z]};
    vec3f P;
    const float d = detail::test(tri, queryPoint, P);
    if (d < result.sqrDist || (d == result.sqrDist && result.primID >= 0 && (int)primID < result.primID)) {
      result.sqrDist = d;
      result.primID = (int)primID;
      result.P = P;
    }
    return result.sqrDist;
  };
  shrinkingRadiusQuery::forEachPrim(perPrim, bvh, queryPoint, result.sqrDist);
}

}  // namespace triangles
}  // namespace cuBQL
