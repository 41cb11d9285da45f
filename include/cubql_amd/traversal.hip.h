// traversal.hip.h -- device-side shrinking-radius traversal for HIP code.
//
// HIP counterpart of cuBQL::shrinkingRadiusQuery::forEachLeaf / forEachPrim
// (traversal/shrinkingRadiusQuery.h:409-508, 766-801) for the BinaryBVH
// float3 layout.  Include from .hip sources compiled for gfx950.
//
// Semantics:
//   * the query walks the tree from node 0, ordering children by their
//     squared distance (ties take the second child first, :463-471);
//   * inclusive culling: a node is visited iff its distance is <= the
//     current cull radius (the reference culls at >=).  This is what the
//     lowest-primID tie rule needs; with it a leaf lambda may see more
//     leaves than in the reference, never fewer;
//   * a leaf lambda returns the new squared cull radius (smaller values
//     shrink the query); a negative value terminates the query;
//   * the traversal stack holds STACK entries; if a push would overflow it
//     the function returns false and the caller must rerun the query with a
//     deeper stack (the reference silently truncates, :475-478).  The stack
//     never needs more entries than the tree is deep.
//   * all distances are computed as (x*x+y*y)+z*z without FMA contraction,
//     which keeps node distances <= the distance of every point inside.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cubql_amd/cubql_capi.h"

namespace cubql_dev {

struct QNode {
  float4 a;   // lo.x lo.y lo.z hi.x
  float4 b;   // hi.y hi.z (oc lo) (oc hi) as bits
};

__device__ __forceinline__ uint64_t node_oc(const QNode &n) {
  return (uint64_t)__float_as_uint(n.b.z) | ((uint64_t)__float_as_uint(n.b.w) << 32);
}
__device__ __forceinline__ uint32_t oc_cnt(uint64_t oc) { return (uint32_t)(oc >> 48); }
__device__ __forceinline__ uint32_t oc_off(uint64_t oc) { return (uint32_t)(oc & ((1ull << 48) - 1)); }

__device__ __forceinline__ QNode load_node(const cubql_node3f *nodes, uint32_t i) {
  const float4 *p = reinterpret_cast<const float4 *>(nodes + i);
  QNode n;
  n.a = p[0];
  n.b = p[1];
  return n;
}

// fSqrDistance_rd(vec3f, box3f) (shrinkingRadiusQuery.h:253-259): clamp to the
// box, then dot without contraction.
__device__ __forceinline__ float sqr_dist_box(float qx, float qy, float qz, const QNode &n) {
#pragma clang fp contract(off)
  float px = fminf(fmaxf(qx, n.a.x), n.a.w);
  float py = fminf(fmaxf(qy, n.a.y), n.b.x);
  float pz = fminf(fmaxf(qz, n.a.z), n.b.y);
  float vx = px - qx, vy = py - qy, vz = pz - qz;
  return vx * vx + vy * vy + vz * vz;
}

// fSqrDistance_rd(vec3f, vec3f) (conservativeDistances.h:61-90).
__device__ __forceinline__ float sqr_dist_pt(float ax, float ay, float az, float bx, float by,
                                             float bz) {
#pragma clang fp contract(off)
  float dx = ax - bx, dy = ay - by, dz = az - bz;
  return dx * dx + dy * dy + dz * dz;
}

// forEachLeaf.  leaf(const uint32_t *prims, uint32_t count) -> float.
// Returns false if the stack overflowed (query incomplete).
// `root` selects the subtree to walk (0 for a whole BinaryBVH; a cloud's
// root node in a batched forest).
template <int STACK, typename LeafFn>
__device__ __forceinline__ bool for_each_leaf(const cubql_bvh3f &bvh, float qx, float qy, float qz,
                                              float sqrMax, LeafFn &&leaf, uint32_t root = 0) {
  if (bvh.numNodes == 0) return true;
  uint2 stack[STACK];
  int sp = 0;
  float cull = sqrMax;
  uint64_t node = node_oc(load_node(bvh.nodes, root));
  while (true) {
    bool haveLeaf = true;
    while (oc_cnt(node) == 0) {
      const uint32_t i0 = oc_off(node);
      const QNode n0 = load_node(bvh.nodes, i0);
      const QNode n1 = load_node(bvh.nodes, i0 + 1);
      const float d0 = sqr_dist_box(qx, qy, qz, n0);
      const float d1 = sqr_dist_box(qx, qy, qz, n1);
      if (fminf(d0, d1) > cull) { haveLeaf = false; break; }
      const bool near0 = d0 < d1;
      node = near0 ? node_oc(n0) : node_oc(n1);
      const float dfar = fmaxf(d0, d1);
      if (dfar <= cull) {
        if (sp >= STACK) return false;
        stack[sp++] = make_uint2(near0 ? i0 + 1 : i0, __float_as_uint(dfar));
      }
    }
    if (haveLeaf) {
      float r = leaf(bvh.primIDs + oc_off(node), oc_cnt(node));
      if (r < 0.f) return true;
      cull = fminf(cull, r);
    }
    bool found = false;
    while (sp > 0) {
      uint2 e = stack[--sp];
      if (__uint_as_float(e.y) <= cull) {
        node = node_oc(load_node(bvh.nodes, e.x));
        found = true;
        break;
      }
    }
    if (!found) return true;
  }
}

// Descends from `root` along the nearer child (ties to the second child, as
// the traversal orders them) to one leaf and calls prim(primID) for its
// prims.  A closest-point query that runs this first can start the full
// traversal with the radius it found, so that on the way down it stacks only
// far children that can still hold a closer prim.
template <typename PrimFn>
__device__ __forceinline__ void visit_nearest_leaf(const cubql_bvh3f &bvh, float qx, float qy, float qz,
                                                   PrimFn &&prim, uint32_t root = 0) {
  if (bvh.numNodes == 0) return;
  uint64_t node = node_oc(load_node(bvh.nodes, root));
  while (oc_cnt(node) == 0) {
    const uint32_t i0 = oc_off(node);
    const QNode n0 = load_node(bvh.nodes, i0);
    const QNode n1 = load_node(bvh.nodes, i0 + 1);
    node = sqr_dist_box(qx, qy, qz, n0) < sqr_dist_box(qx, qy, qz, n1) ? node_oc(n0) : node_oc(n1);
  }
  const uint32_t *ids = bvh.primIDs + oc_off(node);
  for (uint32_t j = 0; j < oc_cnt(node); j++) prim(ids[j]);
}

// forEachPrim: prim(uint32_t primID) -> float (new squared cull radius).
template <int STACK, typename PrimFn>
__device__ __forceinline__ bool for_each_prim(const cubql_bvh3f &bvh, float qx, float qy, float qz,
                                              float sqrMax, PrimFn &&prim, uint32_t root = 0) {
  auto leaf = [&](const uint32_t *ids, uint32_t cnt) -> float {
    float r = INFINITY;
    for (uint32_t i = 0; i < cnt; i++) r = fminf(r, prim(ids[i]));
    return r;
  };
  return for_each_leaf<STACK>(bvh, qx, qy, qz, sqrMax, leaf, root);
}

}  // namespace cubql_dev
