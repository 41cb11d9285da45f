// cuBQL/traversal/shrinkingRadiusQuery.h -- device-side shrinking-radius
// traversal in the reference's spelling (traversal/shrinkingRadiusQuery.h:
// 55-228 interface, 409-508 and 653-801 implementation) for bvh3f, and for
// WideBVH<float,3,W> (:553-636).
//
//   forEachLeaf(leafLambda, bvh, queryPoint, sqrMaxSearchRadius = INF)
//     leafLambda(const uint32_t *primIDs, int count) -> float
//   forEachPrim(primLambda, bvh, queryPoint, sqrMaxSearchRadius = INF)
//     primLambda(uint32_t primID) -> float
//   forEachLeaf / forEachPrim(lambda, nodeLambda, bvh, sqrMaxSearchRadius)
//     nodeLambda(const bvh3f::Node &) -> float, a conservative (never
//     larger than true) squared distance to the node's contents
//
// A lambda returns the squared radius the query may shrink to (larger
// values are ignored); a negative return ends the query.  Rules, shared
// with the library's batched queries so that results agree with them: the
// nearer child is visited first (equal distances: the second child first);
// a node is visited iff its squared distance is <= the current radius
// (inclusive; the reference culls at >=, which leaves the choice among
// equal-distance results to the traversal order); distances are
// (x*x + y*y) + z*z without FMA contraction.  The traversal stack holds
// CUBQL_TRAVERSAL_STACK entries (default 64, as in the reference); a query
// that would push beyond it stops there, as the reference's does
// (shrinkingRadiusQuery.h:475-478).  The library's batched queries never
// stop early: they rerun such queries with a deeper stack.
#pragma once
#include "cuBQL/bvh.h"
#include "cubql_amd/traversal.hip.h"

#ifndef CUBQL_TRAVERSAL_STACK
#define CUBQL_TRAVERSAL_STACK 64
#endif

namespace cuBQL {
namespace shrinkingRadiusQuery {

namespace detail {
// One traversal over a binary BVH with a caller-supplied node distance.
template <int STACK, typename NodeDist, typename LeafFn>
__device__ inline void traverse(const bvh3f &bvh, float sqrMax, const NodeDist &nodeDist,
                                const LeafFn &leaf) {
  if (bvh.numNodes == 0) return;
  uint2 stack[STACK];
  int sp = 0;
  float cull = sqrMax;
  bvh3f::Node::Admin node = bvh.nodes[0].admin;
  while (true) {
    bool haveLeaf = true;
    while (node.count == 0) {
      const uint32_t i0 = (uint32_t)node.offset;
      const bvh3f::Node n0 = bvh.nodes[i0];
      const bvh3f::Node n1 = bvh.nodes[i0 + 1];
      const float d0 = nodeDist(n0);
      const float d1 = nodeDist(n1);
      if (fminf(d0, d1) > cull) { haveLeaf = false; break; }
      const bool near0 = d0 < d1;
      node = near0 ? n0.admin : n1.admin;
      const float dfar = fmaxf(d0, d1);
      if (dfar <= cull) {
        if (sp >= STACK) return;
        stack[sp++] = make_uint2(near0 ? i0 + 1 : i0, __float_as_uint(dfar));
      }
    }
    if (haveLeaf) {
      const float r = leaf(bvh.primIDs + node.offset, (int)node.count);
      if (r < 0.f) return;
      cull = fminf(cull, r);
    }
    bool found = false;
    while (sp > 0) {
      const uint2 e = stack[--sp];
      if (__uint_as_float(e.y) <= cull) {
        node = bvh.nodes[e.x].admin;
        found = true;
        break;
      }
    }
    if (!found) return;
  }
}
}  // namespace detail

template <typename Lambda>
__device__ inline void forEachLeaf(const Lambda &lambdaToExecuteForEachCandidateLeaf, bvh3f bvh,
                                   vec3f queryPoint, float sqrMaxSearchRadius = INFINITY) {
  detail::traverse<CUBQL_TRAVERSAL_STACK>(
      bvh, sqrMaxSearchRadius,
      [&](const bvh3f::Node &n) { return fSqrDistance_rd(queryPoint, n.bounds); },
      lambdaToExecuteForEachCandidateLeaf);
}

template <typename Lambda>
__device__ inline void forEachPrim(const Lambda &lambdaToExecuteForEachCandidate, bvh3f bvh,
                                   vec3f queryPoint, float sqrMaxSearchRadius = INFINITY) {
  auto leaf = [&](const uint32_t *ids, int cnt) -> float {
    float r = INFINITY;
    for (int i = 0; i < cnt; i++) r = fminf(r, lambdaToExecuteForEachCandidate(ids[i]));
    return r;
  };
  forEachLeaf(leaf, bvh, queryPoint, sqrMaxSearchRadius);
}

template <typename LeafLambda, typename NodeLambda>
__device__ inline void forEachLeaf(const LeafLambda &lambdaToExecuteForEachCandidateLeaf,
                                   const NodeLambda &lambdaToExecuteForEachCandidateNode, bvh3f bvh,
                                   float sqrMaxSearchRadius = INFINITY) {
  detail::traverse<CUBQL_TRAVERSAL_STACK>(bvh, sqrMaxSearchRadius, lambdaToExecuteForEachCandidateNode,
                                          lambdaToExecuteForEachCandidateLeaf);
}

template <typename PrimLambda, typename NodeLambda>
__device__ inline void forEachPrim(const PrimLambda &lambdaToExecuteForEachCandidatePrim,
                                   const NodeLambda &lambdaToExecuteForEachVisitedNode, bvh3f bvh,
                                   float sqrMaxSearchRadius = INFINITY) {
  auto leaf = [&](const uint32_t *ids, int cnt) -> float {
    float r = INFINITY;
    for (int i = 0; i < cnt; i++) r = fminf(r, lambdaToExecuteForEachCandidatePrim(ids[i]));
    return r;
  };
  forEachLeaf(leaf, lambdaToExecuteForEachVisitedNode, bvh, sqrMaxSearchRadius);
}

/*! forEachLeaf / forEachPrim over a WideBVH<float,3,W> (W = 4, 8;
    shrinkingRadiusQuery.h:553-636): children are visited nearest first
    under the same inclusive rule as the binary traversal. */
template <int W, typename Lambda>
__device__ inline void forEachLeaf(const Lambda &lambdaToExecuteForEachCandidateLeaf, WideBVH<float, 3, W> bvh,
                                   vec3f queryPoint, float sqrMaxSearchRadius = INFINITY) {
  const cubql_dev::WideBVH<W> v{reinterpret_cast<const float4 *>(bvh.nodes), bvh.numNodes, bvh.primIDs,
                                bvh.numPrims};
  auto leaf = [&](const uint32_t *ids, uint32_t cnt) -> float {
    return lambdaToExecuteForEachCandidateLeaf(ids, (int)cnt);
  };
  cubql_dev::for_each_leaf<CUBQL_TRAVERSAL_STACK>(v, queryPoint.x, queryPoint.y, queryPoint.z, sqrMaxSearchRadius,
                                                  leaf);
}

template <int W, typename Lambda>
__device__ inline void forEachPrim(const Lambda &lambdaToExecuteForEachCandidate, WideBVH<float, 3, W> bvh,
                                   vec3f queryPoint, float sqrMaxSearchRadius = INFINITY) {
  auto leaf = [&](const uint32_t *ids, int cnt) -> float {
    float r = INFINITY;
    for (int i = 0; i < cnt; i++) r = fminf(r, lambdaToExecuteForEachCandidate(ids[i]));
    return r;
  };
  forEachLeaf(leaf, bvh, queryPoint, sqrMaxSearchRadius);
}

}  // namespace shrinkingRadiusQuery
}  // namespace cuBQL
