// cuBQL/queries/points/knn.h -- device k-nearest-neighbour query on a bvh3f
// built over points (queries/points/knn.h:49-116).
//
//   knn::Result findKNN(knn::Candidate *foundCandidates, int maxN,
//                       bvh3f bvh, const vec3f *points, vec3f queryPoint,
//                       float squareOfMaxQueryDistance = INFINITY)
//
// foundCandidates receives up to maxN candidates sorted by (sqrDist,
// primID) (cuBQL/queries/common/knn.h); the result equals a row of the
// library's batched cubql_knn_points_f3 for the same tree and query.
#pragma once
#include "cuBQL/queries/common/knn.h"
#include "cuBQL/traversal/shrinkingRadiusQuery.h"

namespace cuBQL {
namespace points {

__device__ inline knn::Result findKNN(knn::Candidate *foundCandidates, int maxN, bvh3f bvhOverPoints,
                                      const vec3f *points, vec3f queryPoint,
                                      float squareOfMaxQueryDistance = INFINITY) {
  knn::Result result = {0, squareOfMaxQueryDistance};
  auto candidateLambda = [&](uint32_t pointID) -> float {
    const float sqrDist = fSqrDistance_rd(points[pointID], queryPoint);
    if (!(sqrDist < squareOfMaxQueryDistance)) return result.sqrDistMax;
    knn::insert_linear(foundCandidates, maxN, knn::Candidate{(int)pointID, sqrDist}, result);
    return result.sqrDistMax;
  };
  shrinkingRadiusQuery::forEachPrim(candidateLambda, bvhOverPoints, queryPoint, squareOfMaxQueryDistance);
  return result;
}

}  // namespace points
}  // namespace cuBQL
