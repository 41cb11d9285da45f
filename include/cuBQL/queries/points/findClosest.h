// cuBQL/queries/points/findClosest.h -- device closest-point query on a
// bvh3f built over points (queries/points/findClosest.h:49-62, 181-226).
//
//   int findClosest(bvh3f bvh, const vec3f *points, vec3f queryPoint,
//                   float squareOfMaxQueryDistance = INFINITY)
//
// Returns the index of the closest point within the radius, or -1.  Among
// equal squared distances the lowest index is returned (the reference
// leaves that choice to the traversal order), so the result equals the
// library's batched cubql_fcp_points_f3 for the same tree and query.  The
// query first descends to the nearest-child leaf to seed the radius, then
// runs the shrinking-radius traversal.
#pragma once
#include "cuBQL/traversal/shrinkingRadiusQuery.h"

namespace cuBQL {
namespace points {

namespace detail {
// descend along the nearer child (second child on ties) to one leaf
template <typename PrimFn>
__device__ inline void visitNearestLeaf(const bvh3f &bvh, vec3f q, const PrimFn &prim) {
  if (bvh.numNodes == 0) return;
  bvh3f::Node::Admin node = bvh.nodes[0].admin;
  while (node.count == 0) {
    const bvh3f::Node &n0 = bvh.nodes[node.offset];
    const bvh3f::Node &n1 = bvh.nodes[node.offset + 1];
    node = fSqrDistance_rd(q, n0.bounds) < fSqrDistance_rd(q, n1.bounds) ? n0.admin : n1.admin;
  }
  for (uint32_t j = 0; j < node.count; j++) prim(bvh.primIDs[node.offset + j]);
}
}  // namespace detail

__device__ inline int findClosest(bvh3f bvhOverPoints, const vec3f *points, vec3f queryPoint,
                                  float squareOfMaxQueryDistance = INFINITY) {
  int best = -1;
  float bestD = squareOfMaxQueryDistance;
  auto prim = [&](uint32_t pid) -> float {
    const float d = fSqrDistance_rd(points[pid], queryPoint);
    if (d < bestD || (d == bestD && best >= 0 && (int)pid < best)) {
      bestD = d;
      best = (int)pid;
    }
    return bestD;
  };
  detail::visitNearestLeaf(bvhOverPoints, queryPoint, prim);
  shrinkingRadiusQuery::forEachPrim(prim, bvhOverPoints, queryPoint, bestD);
  return best;
}

/*! findClosest over a WideBVH<float,3,W> (findClosest.h:82-90, 228-258) */
template <int W>
__device__ inline int findClosest(WideBVH<float, 3, W> bvhOverPoints, const vec3f *points, vec3f queryPoint,
                                  float squareOfMaxQueryDistance = INFINITY) {
  int best = -1;
  float bestD = squareOfMaxQueryDistance;
  auto prim = [&](uint32_t pid) -> float {
    const float d = fSqrDistance_rd(points[pid], queryPoint);
    if (d < bestD || (d == bestD && best >= 0 && (int)pid < best)) {
      bestD = d;
      best = (int)pid;
    }
    return bestD;
  };
  shrinkingRadiusQuery::forEachPrim(prim, bvhOverPoints, queryPoint, bestD);
  return best;
}

/*! findClosest_exludeID (findClosest.h:95-110): the same, ignoring one
    point index */
__device__ inline int findClosest_exludeID(int idOfPointtoExclude, bvh3f bvhOverPoints, const vec3f *points,
                                           vec3f queryPoint, float squareOfMaxQueryDistance = INFINITY) {
  int best = -1;
  float bestD = squareOfMaxQueryDistance;
  auto prim = [&](uint32_t pid) -> float {
    if ((int)pid == idOfPointtoExclude) return bestD;
    const float d = fSqrDistance_rd(points[pid], queryPoint);
    if (d < bestD || (d == bestD && best >= 0 && (int)pid < best)) {
      bestD = d;
      best = (int)pid;
    }
    return bestD;
  };
  shrinkingRadiusQuery::forEachPrim(prim, bvhOverPoints, queryPoint, bestD);
  return best;
}

/*! findClosest_exludeSelf (findClosest.h:112-126): ignores data points
    at the query position itself */
__device__ inline int findClosest_exludeSelf(bvh3f bvhOverPoints, const vec3f *points, vec3f queryPoint,
                                             float squareOfMaxQueryDistance = INFINITY) {
  int best = -1;
  float bestD = squareOfMaxQueryDistance;
  auto prim = [&](uint32_t pid) -> float {
    const vec3f p = points[pid];
    if (p == queryPoint) return bestD;
    const float d = fSqrDistance_rd(p, queryPoint);
    if (d < bestD || (d == bestD && best >= 0 && (int)pid < best)) {
      bestD = d;
      best = (int)pid;
    }
    return bestD;
  };
  shrinkingRadiusQuery::forEachPrim(prim, bvhOverPoints, queryPoint, bestD);
  return best;
}

}  // namespace points
}  // namespace cuBQL
