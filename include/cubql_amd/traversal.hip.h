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

// Traversal counters for the instrumented query kernels (the counterpart of
// the nodeDist / primCode counters of testing/bvh/buildQual.h:65-98): bytes
// of node data loaded (a 64-byte child pair per inner node expanded, a
// 32-byte node per stack entry popped, W 32-byte children per wide node)
// and prims tested (counted by the query's prim code).  NoStats compiles to
// nothing.
struct NoStats {
  __device__ __forceinline__ void nodes(uint32_t) {}
  __device__ __forceinline__ void prim() {}
};
struct TraversalCounts {
  uint32_t nodeBytes = 0, prims = 0;
  __device__ __forceinline__ void nodes(uint32_t b) { nodeBytes += b; }
  __device__ __forceinline__ void prim() { prims++; }
};

// Traversal stack storage for (node id, distance bits) entries; the
// traversal keeps the stack pointer itself (in a register).  PrivateStack
// is an array in the lane's private memory (the compiler places it in
// scratch); LdsStack keeps N entries per lane in LDS, entry j of a lane at
// base[j * STRIDE] (base = the workgroup's array + the lane's thread index,
// STRIDE = the workgroup size), so a wave's accesses to one depth fall in
// consecutive banks.
template <int N>
struct PrivateStack {
  static constexpr int kCap = N;
  uint2 e[N];
  __device__ __forceinline__ void put(int i, uint2 v) { e[i] = v; }
  __device__ __forceinline__ uint2 get(int i) const { return e[i]; }
};
template <int N, int STRIDE>
struct LdsStack {
  static constexpr int kCap = N;
  uint2 *base;
  __device__ __forceinline__ explicit LdsStack(uint2 *b) : base(b) {}
  __device__ __forceinline__ void put(int i, uint2 v) { base[i * STRIDE] = v; }
  __device__ __forceinline__ uint2 get(int i) const { return base[i * STRIDE]; }
};

// forEachLeaf.  leaf(const uint32_t *prims, uint32_t count) -> float.
// Returns false if the stack overflowed (query incomplete).
// `root` selects the subtree to walk (0 for a whole BinaryBVH; a cloud's
// root node in a batched forest).
template <typename StackT, typename LeafFn, typename Stats>
__device__ __forceinline__ bool for_each_leaf_on(StackT &stack, const cubql_bvh3f &bvh, float qx, float qy,
                                                 float qz, float sqrMax, LeafFn &&leaf, uint32_t root, Stats &st) {
  if (bvh.numNodes == 0) return true;
  int sp = 0;
  float cull = sqrMax;
  uint64_t node = node_oc(load_node(bvh.nodes, root));
  while (true) {
    bool haveLeaf = true;
    while (oc_cnt(node) == 0) {
      const uint32_t i0 = oc_off(node);
      const QNode n0 = load_node(bvh.nodes, i0);
      const QNode n1 = load_node(bvh.nodes, i0 + 1);
      st.nodes(64);
      const float d0 = sqr_dist_box(qx, qy, qz, n0);
      const float d1 = sqr_dist_box(qx, qy, qz, n1);
      if (fminf(d0, d1) > cull) { haveLeaf = false; break; }
      const bool near0 = d0 < d1;
      node = near0 ? node_oc(n0) : node_oc(n1);
      const float dfar = fmaxf(d0, d1);
      if (dfar <= cull) {
        if (sp >= StackT::kCap) return false;
        stack.put(sp++, make_uint2(near0 ? i0 + 1 : i0, __float_as_uint(dfar)));
      }
    }
    if (haveLeaf) {
      float r = leaf(bvh.primIDs + oc_off(node), oc_cnt(node));
      if (r < 0.f) return true;
      cull = fminf(cull, r);
    }
    bool found = false;
    while (sp > 0) {
      uint2 e = stack.get(--sp);
      if (__uint_as_float(e.y) <= cull) {
        node = node_oc(load_node(bvh.nodes, e.x));
        st.nodes(32);
        found = true;
        break;
      }
    }
    if (!found) return true;
  }
}
template <int STACK, typename LeafFn, typename Stats>
__device__ __forceinline__ bool for_each_leaf(const cubql_bvh3f &bvh, float qx, float qy, float qz,
                                              float sqrMax, LeafFn &&leaf, uint32_t root, Stats &st) {
  PrivateStack<STACK> stack;
  return for_each_leaf_on(stack, bvh, qx, qy, qz, sqrMax, leaf, root, st);
}
template <int STACK, typename LeafFn>
__device__ __forceinline__ bool for_each_leaf(const cubql_bvh3f &bvh, float qx, float qy, float qz,
                                              float sqrMax, LeafFn &&leaf, uint32_t root = 0) {
  NoStats st;
  return for_each_leaf<STACK>(bvh, qx, qy, qz, sqrMax, leaf, root, st);
}

// Descends from `root` along the nearer child (ties to the second child, as
// the traversal orders them) to one leaf and calls prim(primID) for its
// prims.  A closest-point query that runs this first can start the full
// traversal with the radius it found, so that on the way down it stacks only
// far children that can still hold a closer prim.  Returns the leaf's first
// primID slot (kNoLeaf for an empty tree), which for_each_prim_except
// leaves out.
constexpr uint32_t kNoLeaf = 0xffffffffu;
// The offset/count word of the leaf that descent ends in (the tree must not
// be empty).
template <typename Stats>
__device__ __forceinline__ uint64_t nearest_leaf(const cubql_bvh3f &bvh, float qx, float qy, float qz,
                                                 uint32_t root, Stats &st) {
  uint64_t node = node_oc(load_node(bvh.nodes, root));
  while (oc_cnt(node) == 0) {
    const uint32_t i0 = oc_off(node);
    const QNode n0 = load_node(bvh.nodes, i0);
    const QNode n1 = load_node(bvh.nodes, i0 + 1);
    st.nodes(64);
    node = sqr_dist_box(qx, qy, qz, n0) < sqr_dist_box(qx, qy, qz, n1) ? node_oc(n0) : node_oc(n1);
  }
  return node;
}
template <typename PrimFn, typename Stats>
__device__ __forceinline__ uint32_t visit_nearest_leaf(const cubql_bvh3f &bvh, float qx, float qy, float qz,
                                                       PrimFn &&prim, uint32_t root, Stats &st) {
  if (bvh.numNodes == 0) return kNoLeaf;
  const uint64_t node = nearest_leaf(bvh, qx, qy, qz, root, st);
  const uint32_t *ids = bvh.primIDs + oc_off(node);
  for (uint32_t j = 0; j < oc_cnt(node); j++) prim(ids[j]);
  return oc_off(node);
}
template <typename PrimFn>
__device__ __forceinline__ uint32_t visit_nearest_leaf(const cubql_bvh3f &bvh, float qx, float qy, float qz,
                                                       PrimFn &&prim, uint32_t root = 0) {
  NoStats st;
  return visit_nearest_leaf(bvh, qx, qy, qz, prim, root, st);
}

// forEachPrim: prim(uint32_t primID) -> float (new squared cull radius).
template <int STACK, typename PrimFn, typename Stats>
__device__ __forceinline__ bool for_each_prim(const cubql_bvh3f &bvh, float qx, float qy, float qz,
                                              float sqrMax, PrimFn &&prim, uint32_t root, Stats &st) {
  auto leaf = [&](const uint32_t *ids, uint32_t cnt) -> float {
    float r = INFINITY;
    for (uint32_t i = 0; i < cnt; i++) r = fminf(r, prim(ids[i]));
    return r;
  };
  return for_each_leaf<STACK>(bvh, qx, qy, qz, sqrMax, leaf, root, st);
}
template <int STACK, typename PrimFn>
__device__ __forceinline__ bool for_each_prim(const cubql_bvh3f &bvh, float qx, float qy, float qz,
                                              float sqrMax, PrimFn &&prim, uint32_t root = 0) {
  NoStats st;
  return for_each_prim<STACK>(bvh, qx, qy, qz, sqrMax, prim, root, st);
}
// forEachPrim on a caller-provided stack (PrivateStack / LdsStack).
template <typename StackT, typename PrimFn, typename Stats>
__device__ __forceinline__ bool for_each_prim_on(StackT &stack, const cubql_bvh3f &bvh, float qx, float qy,
                                                 float qz, float sqrMax, PrimFn &&prim, uint32_t root, Stats &st) {
  auto leaf = [&](const uint32_t *ids, uint32_t cnt) -> float {
    float r = INFINITY;
    for (uint32_t i = 0; i < cnt; i++) r = fminf(r, prim(ids[i]));
    return r;
  };
  return for_each_leaf_on(stack, bvh, qx, qy, qz, sqrMax, leaf, root, st);
}

// forEachPrim without the leaf whose first primID slot is `skip` (the leaf
// visit_nearest_leaf returned: the caller has tested its prims, and testing
// a prim again changes no closest-point result).  A left-out leaf leaves the
// radius as it is.
template <typename PrimFn>
struct PrimsExcept {
  const uint32_t *primIDs;
  uint32_t skip;
  PrimFn &prim;
  __device__ __forceinline__ float operator()(const uint32_t *ids, uint32_t cnt) const {
    if ((uint32_t)(ids - primIDs) == skip) return INFINITY;
    float r = INFINITY;
    for (uint32_t i = 0; i < cnt; i++) r = fminf(r, prim(ids[i]));
    return r;
  }
};
template <typename StackT, typename PrimFn, typename Stats>
__device__ __forceinline__ bool for_each_prim_except_on(StackT &stack, const cubql_bvh3f &bvh, float qx, float qy,
                                                        float qz, float sqrMax, PrimFn &&prim, uint32_t skip,
                                                        uint32_t root, Stats &st) {
  PrimsExcept<PrimFn> leaf{bvh.primIDs, skip, prim};
  return for_each_leaf_on(stack, bvh, qx, qy, qz, sqrMax, leaf, root, st);
}
template <int STACK, typename PrimFn, typename Stats>
__device__ __forceinline__ bool for_each_prim_except(const cubql_bvh3f &bvh, float qx, float qy, float qz,
                                                     float sqrMax, PrimFn &&prim, uint32_t skip, uint32_t root,
                                                     Stats &st) {
  PrivateStack<STACK> stack;
  return for_each_prim_except_on(stack, bvh, qx, qy, qz, sqrMax, prim, skip, root, st);
}

}  // namespace cubql_dev

// ------------------------------------------------------------ wide BVHs
//
// WideBVH<float,3,W> (bvh.h:113-141): node i holds W consecutive 32-byte
// children {box, bits}: valid in bit 0, offset in bits 1-45, count in bits
// 46-61.  A valid child with count 0 is an inner child whose wide node is
// nodes[offset]; otherwise a leaf over primIDs[offset, offset+count).
namespace cubql_dev {

template <int W>
struct WideBVH {
  const float4 *nodes;  // numNodes * W children, two float4 each
  uint32_t numNodes;
  const uint32_t *primIDs;
  uint32_t numPrims;
};

__device__ __forceinline__ bool wc_valid(uint64_t w) { return (w & 1ull) != 0; }
__device__ __forceinline__ uint32_t wc_off(uint64_t w) { return (uint32_t)((w >> 1) & ((1ull << 45) - 1)); }
__device__ __forceinline__ uint32_t wc_cnt(uint64_t w) { return (uint32_t)((w >> 46) & 0xffffu); }

template <int W> struct Log2;
template <> struct Log2<4> { static constexpr int value = 2; };
template <> struct Log2<8> { static constexpr int value = 3; };

// Children of wide node `node` in visiting order: key = (squared distance
// bits << 32) | payload, ascending, with ~0 for invalid children and
// children farther than `cull`.  Payload: the wide node id of an inner
// child, (1 << 31) | (node << log2 W) | slot for a leaf child (the
// reference's ChildOrder, shrinkingRadiusQuery.h:511-537).
template <int W>
__device__ __forceinline__ void wide_child_keys(const WideBVH<W> &bvh, uint32_t node, float qx, float qy, float qz,
                                                float cull, uint64_t (&key)[W]) {
  const float4 *c = bvh.nodes + (size_t)node * W * 2;
#pragma unroll
  for (int j = 0; j < W; j++) {
    const float4 a = c[2 * j], b = c[2 * j + 1];
    const uint64_t bits = (uint64_t)__float_as_uint(b.z) | ((uint64_t)__float_as_uint(b.w) << 32);
    QNode n;
    n.a = a;
    n.b = b;
    const float d = sqr_dist_box(qx, qy, qz, n);
    const uint32_t payload = wc_cnt(bits) ? (0x80000000u | (node << Log2<W>::value) | (uint32_t)j) : wc_off(bits);
    key[j] = (wc_valid(bits) && d <= cull) ? (((uint64_t)__float_as_uint(d) << 32) | payload) : ~0ull;
  }
#pragma unroll
  for (int i = W - 1; i > 0; --i) {
#pragma unroll
    for (int j = 0; j < i; j++) {
      const uint64_t k0 = key[j], k1 = key[j + 1];
      key[j] = k0 < k1 ? k0 : k1;
      key[j + 1] = k0 < k1 ? k1 : k0;
    }
  }
}

template <int W>
__device__ __forceinline__ uint64_t wide_leaf_bits(const WideBVH<W> &bvh, uint32_t payload, int log2w) {
  const uint32_t node = (payload & 0x7fffffffu) >> log2w, slot = payload & ((1u << log2w) - 1);
  const float4 b = bvh.nodes[((size_t)node << log2w) * 2 + 2 * slot + 1];
  return (uint64_t)__float_as_uint(b.z) | ((uint64_t)__float_as_uint(b.w) << 32);
}

// forEachLeaf over a wide BVH (shrinkingRadiusQuery.h:553-636) with the
// binary traversal's rules: children visited nearest first, a child is
// visited iff its distance is <= the current radius, a negative leaf
// result ends the query.  Returns false if the stack overflowed.
template <int STACK, int W, typename LeafFn, typename Stats>
__device__ __forceinline__ bool for_each_leaf(const WideBVH<W> &bvh, float qx, float qy, float qz, float sqrMax,
                                              LeafFn &&leaf, uint32_t root, Stats &st) {
  if (bvh.numNodes == 0) return true;
  uint64_t stack[STACK];
  int sp = 0;
  float cull = sqrMax;
  uint32_t node = root;
  bool expand = true;
  while (true) {
    if (expand) {
      uint64_t key[W];
      wide_child_keys<W>(bvh, node, qx, qy, qz, cull, key);
      st.nodes(32 * W);
#pragma unroll
      for (int j = W - 1; j > 0; --j) {
        if (key[j] != ~0ull) {
          if (sp >= STACK) return false;
          stack[sp++] = key[j];
        }
      }
      expand = false;
      if (key[0] != ~0ull) {
        const uint32_t payload = (uint32_t)key[0];
        if (payload & 0x80000000u) {
          const uint64_t bits = wide_leaf_bits(bvh, payload, Log2<W>::value);
          const float r = leaf(bvh.primIDs + wc_off(bits), wc_cnt(bits));
          if (r < 0.f) return true;
          cull = fminf(cull, r);
        } else {
          node = payload;
          expand = true;
          continue;
        }
      }
    }
    // pop the next entry still within the radius
    bool found = false;
    while (sp > 0) {
      const uint64_t e = stack[--sp];
      if (__uint_as_float((uint32_t)(e >> 32)) > cull) continue;
      const uint32_t payload = (uint32_t)e;
      if (payload & 0x80000000u) {
        const uint64_t bits = wide_leaf_bits(bvh, payload, Log2<W>::value);
        const float r = leaf(bvh.primIDs + wc_off(bits), wc_cnt(bits));
        if (r < 0.f) return true;
        cull = fminf(cull, r);
        continue;
      }
      node = payload;
      found = true;
      break;
    }
    if (!found) return true;
    expand = true;
  }
}
template <int STACK, int W, typename LeafFn>
__device__ __forceinline__ bool for_each_leaf(const WideBVH<W> &bvh, float qx, float qy, float qz, float sqrMax,
                                              LeafFn &&leaf, uint32_t root = 0) {
  NoStats st;
  return for_each_leaf<STACK>(bvh, qx, qy, qz, sqrMax, leaf, root, st);
}

// Descends along the nearest child to one leaf and calls prim(primID) for
// its prims (the wide counterpart of visit_nearest_leaf).
template <int W, typename PrimFn, typename Stats>
__device__ __forceinline__ uint32_t visit_nearest_leaf(const WideBVH<W> &bvh, float qx, float qy, float qz,
                                                       PrimFn &&prim, uint32_t root, Stats &st) {
  if (bvh.numNodes == 0) return kNoLeaf;
  uint32_t node = root;
  for (int guard = 0; guard < (1 << 20); guard++) {
    uint64_t key[W];
    wide_child_keys<W>(bvh, node, qx, qy, qz, INFINITY, key);
    st.nodes(32 * W);
    if (key[0] == ~0ull) return kNoLeaf;
    const uint32_t payload = (uint32_t)key[0];
    if (payload & 0x80000000u) {
      const uint64_t bits = wide_leaf_bits(bvh, payload, Log2<W>::value);
      const uint32_t *ids = bvh.primIDs + wc_off(bits);
      for (uint32_t j = 0; j < wc_cnt(bits); j++) prim(ids[j]);
      return wc_off(bits);
    }
    node = payload;
  }
  return kNoLeaf;
}
template <int W, typename PrimFn>
__device__ __forceinline__ uint32_t visit_nearest_leaf(const WideBVH<W> &bvh, float qx, float qy, float qz,
                                                       PrimFn &&prim, uint32_t root = 0) {
  NoStats st;
  return visit_nearest_leaf(bvh, qx, qy, qz, prim, root, st);
}

template <int STACK, int W, typename PrimFn, typename Stats>
__device__ __forceinline__ bool for_each_prim(const WideBVH<W> &bvh, float qx, float qy, float qz, float sqrMax,
                                              PrimFn &&prim, uint32_t root, Stats &st) {
  auto leaf = [&](const uint32_t *ids, uint32_t cnt) -> float {
    float r = INFINITY;
    for (uint32_t i = 0; i < cnt; i++) r = fminf(r, prim(ids[i]));
    return r;
  };
  return for_each_leaf<STACK>(bvh, qx, qy, qz, sqrMax, leaf, root, st);
}
template <int STACK, int W, typename PrimFn>
__device__ __forceinline__ bool for_each_prim(const WideBVH<W> &bvh, float qx, float qy, float qz, float sqrMax,
                                              PrimFn &&prim, uint32_t root = 0) {
  NoStats st;
  return for_each_prim<STACK>(bvh, qx, qy, qz, sqrMax, prim, root, st);
}
// (the wide counterpart of the binary for_each_prim_except)
template <int STACK, int W, typename PrimFn, typename Stats>
__device__ __forceinline__ bool for_each_prim_except(const WideBVH<W> &bvh, float qx, float qy, float qz,
                                                     float sqrMax, PrimFn &&prim, uint32_t skip, uint32_t root,
                                                     Stats &st) {
  PrimsExcept<PrimFn> leaf{bvh.primIDs, skip, prim};
  return for_each_leaf<STACK>(bvh, qx, qy, qz, sqrMax, leaf, root, st);
}

}  // namespace cubql_dev

// ------------------------------------------------------- fixed box queries
//
// cuBQL::fixedBoxQuery::forEachLeaf (traversal/fixedBoxQuery.h:74-139 for
// binary BVHs, :163-212 for wide ones): every leaf whose box overlaps a
// fixed query box (closed boxes: overlap unless one box's upper is below
// the other's lower on some axis, box.h:208-212), depth first, first child
// first.  leaf(const uint32_t *primIDs, uint32_t count) returns nonzero to
// end the query.  Returns false if the stack overflowed.
namespace cubql_dev {

__device__ __forceinline__ bool box_overlaps(const float (&lo)[3], const float (&hi)[3], const QNode &n) {
  return !(n.a.w < lo[0] || n.b.x < lo[1] || n.b.y < lo[2] || hi[0] < n.a.x || hi[1] < n.a.y ||
           hi[2] < n.a.z);
}

template <int STACK, typename LeafFn>
__device__ __forceinline__ bool for_each_leaf_in_box(const cubql_bvh3f &bvh, const float (&lo)[3],
                                                     const float (&hi)[3], LeafFn &&leaf) {
  if (bvh.numNodes == 0) return true;
  uint32_t stack[STACK];
  int sp = 0;
  uint64_t node = node_oc(load_node(bvh.nodes, 0));
  while (true) {
    bool haveLeaf = true;
    while (oc_cnt(node) == 0) {
      const uint32_t i0 = oc_off(node);
      const QNode n0 = load_node(bvh.nodes, i0);
      const QNode n1 = load_node(bvh.nodes, i0 + 1);
      const bool o0 = box_overlaps(lo, hi, n0), o1 = box_overlaps(lo, hi, n1);
      if (!o0 && !o1) { haveLeaf = false; break; }
      if (o0 && o1) {
        if (sp >= STACK) return false;
        stack[sp++] = i0 + 1;
      }
      node = o0 ? node_oc(n0) : node_oc(n1);
    }
    if (haveLeaf && leaf(bvh.primIDs + oc_off(node), oc_cnt(node))) return true;
    if (sp == 0) return true;
    node = node_oc(load_node(bvh.nodes, stack[--sp]));
  }
}

// wide BVH: children of a node in slot order; a stack entry is (node, next
// slot) as in the reference
template <int STACK, int W, typename LeafFn>
__device__ __forceinline__ bool for_each_leaf_in_box(const WideBVH<W> &bvh, const float (&lo)[3],
                                                     const float (&hi)[3], LeafFn &&leaf) {
  if (bvh.numNodes == 0) return true;
  uint2 stack[STACK];
  int sp = 0;
  uint32_t node = 0, slot = 0;
  while (true) {
    const float4 *c = bvh.nodes + ((size_t)node * W + slot) * 2;
    QNode n;
    n.a = c[0];
    n.b = c[1];
    const uint64_t bits = node_oc(n);
    bool advance = true;
    if (wc_valid(bits) && box_overlaps(lo, hi, n)) {
      if (wc_cnt(bits) == 0) {
        if (slot + 1 < W) {
          if (sp >= STACK) return false;
          stack[sp++] = make_uint2(node, slot + 1);
        }
        node = wc_off(bits);
        slot = 0;
        advance = false;
      } else if (leaf(bvh.primIDs + wc_off(bits), wc_cnt(bits))) {
        return true;
      }
    }
    if (advance) {
      if (slot + 1 < W) {
        slot++;
      } else if (sp > 0) {
        const uint2 e = stack[--sp];
        node = e.x;
        slot = e.y;
      } else {
        return true;
      }
    }
  }
}

}  // namespace cubql_dev
