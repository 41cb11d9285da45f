// cuBQL/traversal/fixedBoxQuery.h -- device-side fixed-box queries in the
// reference's spelling (traversal/fixedBoxQuery.h:20-230) for bvh3f and
// WideBVH<float,3,W>:
//
//   forEachLeaf(leafLambda, bvh, queryBox)
//     leafLambda(const uint32_t *primIDs, size_t count) -> int
//   forEachPrim(primLambda, bvh, queryBox)
//     primLambda(int primID) -> int
//
// visit every leaf / prim of a leaf whose box overlaps queryBox (closed
// boxes), depth first, the first child (lowest slot) first; a lambda
// returning CUBQL_TERMINATE_TRAVERSAL ends the query.  The stack holds
// CUBQL_TRAVERSAL_STACK entries (default 64, as the reference); a query that
// would push beyond it stops there.
#pragma once
#include "cuBQL/bvh.h"
#include "cubql_amd/traversal.hip.h"

#ifndef CUBQL_TERMINATE_TRAVERSAL
#define CUBQL_TERMINATE_TRAVERSAL 1
#define CUBQL_CONTINUE_TRAVERSAL 0
#endif
#ifndef CUBQL_TRAVERSAL_STACK
#define CUBQL_TRAVERSAL_STACK 64
#endif

namespace cuBQL {
namespace fixedBoxQuery {

namespace detail {
__device__ inline void bounds(const box3f &b, float (&lo)[3], float (&hi)[3]) {
  lo[0] = b.lower.x; lo[1] = b.lower.y; lo[2] = b.lower.z;
  hi[0] = b.upper.x; hi[1] = b.upper.y; hi[2] = b.upper.z;
}
template <int W>
__device__ inline cubql_dev::WideBVH<W> view(const WideBVH<float, 3, W> &b) {
  return cubql_dev::WideBVH<W>{reinterpret_cast<const float4 *>(b.nodes), b.numNodes, b.primIDs, b.numPrims};
}
}  // namespace detail

template <typename Lambda>
__device__ inline void forEachLeaf(const Lambda &lambdaToCallOnEachLeaf, const bvh3f bvh, const box3f queryBox) {
  float lo[3], hi[3];
  detail::bounds(queryBox, lo, hi);
  auto leaf = [&](const uint32_t *ids, uint32_t cnt) -> int {
    return lambdaToCallOnEachLeaf(ids, (size_t)cnt) == CUBQL_TERMINATE_TRAVERSAL;
  };
  cubql_dev::for_each_leaf_in_box<CUBQL_TRAVERSAL_STACK>(c_bvh(bvh), lo, hi, leaf);
}

template <int W, typename Lambda>
__device__ inline void forEachLeaf(const Lambda &lambdaToCallOnEachLeaf, const WideBVH<float, 3, W> bvh,
                                   const box3f queryBox) {
  float lo[3], hi[3];
  detail::bounds(queryBox, lo, hi);
  auto leaf = [&](const uint32_t *ids, uint32_t cnt) -> int {
    return lambdaToCallOnEachLeaf(ids, (size_t)cnt) == CUBQL_TERMINATE_TRAVERSAL;
  };
  cubql_dev::for_each_leaf_in_box<CUBQL_TRAVERSAL_STACK>(detail::view(bvh), lo, hi, leaf);
}

template <typename BVH, typename Lambda>
__device__ inline void forEachPrim(const Lambda &lambdaToCallOnEachPrim, const BVH bvh, const box3f queryBox) {
  auto leafCode = [&](const uint32_t *primIDs, size_t numPrims) -> int {
    for (int i = 0; i < (int)numPrims; i++)
      if (lambdaToCallOnEachPrim((int)primIDs[i]) == CUBQL_TERMINATE_TRAVERSAL) return CUBQL_TERMINATE_TRAVERSAL;
    return CUBQL_CONTINUE_TRAVERSAL;
  };
  forEachLeaf(leafCode, bvh, queryBox);
}

}  // namespace fixedBoxQuery
}  // namespace cuBQL
