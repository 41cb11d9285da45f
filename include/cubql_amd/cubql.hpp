// cubql.hpp -- host C++ API in the reference's spelling (namespace cuBQL):
// the builder API of cuBQL/builder/cuda.h plus batched, host-callable forms
// of the point and triangle queries over the C ABI (cubql_capi.h).
//
// Include from code compiled by hipcc (or any C++17 compiler with the ROCm
// include path and __HIP_PLATFORM_AMD__); link libcubql_amd.so and
// libamdhip64.so.  Device-side queries for HIP kernels are in
// cuBQL/queries/... and cuBQL/traversal/....
#pragma once

#include "cuBQL/builder/cuda.h"

namespace cuBQL {

/*! Batched host-callable forms of the device queries: one query per
    element, results in device arrays (see cubql_capi.h for semantics). */
namespace points {
/*! points::findClosest (queries/points/findClosest.h:49-62): id of the
    closest point within sqrt(sqrMaxSearchRadius) (-1 if none), ties to the
    lowest primID; sqrDist may be null. */
inline void findClosest(const bvh3f &bvh, const float *d_points, const float *d_queries, uint32_t numQueries,
                        int32_t *d_ids, float *d_sqrDist = nullptr, float sqrMaxSearchRadius = INFINITY,
                        hipStream_t s = 0) {
  detail::check(cubql_fcp_points_f3(reinterpret_cast<const cubql_bvh3f *>(&bvh), d_points, d_queries, numQueries,
                                    sqrMaxSearchRadius, d_ids, d_sqrDist, s),
                "cuBQL::points::findClosest");
}
/*! points::findKNN (queries/points/knn.h:49-68): k <= 64 nearest per query,
    ordered by (distance, primID); unfilled slots -1 / +inf. */
inline void findKNN(const bvh3f &bvh, const float *d_points, const float *d_queries, uint32_t numQueries, int k,
                    int32_t *d_ids, float *d_sqrDist, int32_t *d_numFound = nullptr,
                    float sqrMaxSearchRadius = INFINITY, hipStream_t s = 0) {
  detail::check(cubql_knn_points_f3(reinterpret_cast<const cubql_bvh3f *>(&bvh), d_points, d_queries, numQueries, k,
                                    sqrMaxSearchRadius, d_ids, d_sqrDist, d_numFound, s),
                "cuBQL::points::findKNN");
}
}  // namespace points

namespace triangles {
/*! triangles::findClosest (queries/triangles/findClosest.h:132-152) for a
    mesh (d_tris = int3[] or null for the non-indexed 3-vertex layout). */
inline void findClosest(const bvh3f &bvh, const int32_t *d_tris, const float *d_verts, const float *d_queries,
                        uint32_t numQueries, int32_t *d_triIDs, float *d_sqrDist, float *d_closest = nullptr,
                        float sqrMaxSearchRadius = INFINITY, hipStream_t s = 0) {
  detail::check(cubql_fcp_triangles_f3(reinterpret_cast<const cubql_bvh3f *>(&bvh), d_tris, d_verts, d_queries,
                                       numQueries, sqrMaxSearchRadius, d_triIDs, d_sqrDist, d_closest, s),
                "cuBQL::triangles::findClosest");
}
}  // namespace triangles

}  // namespace cuBQL
