// cuBQL/bvh.h -- BVH types in the reference's spelling (bvh.h:25-165) for
// the float, 3-D instantiation: BuildConfig, BinaryBVH<float,3> (bvh3f) with
// Node{bounds, admin{offset:48, count:16}}, and WideBVH<float,3,W>.  Byte
// layouts equal the reference's and the C ABI's (cubql_capi.h), so a bvh3f
// built through the C ABI is read by these types and vice versa.
#pragma once
#include "cuBQL/math/vec.h"

namespace cuBQL {

/*! BuildConfig (bvh.h:31-64). */
struct BuildConfig {
  enum BuildMethod { SPATIAL_MEDIAN = 0, SAH, ELH };
  BuildConfig &enableSAH() { buildMethod = SAH; return *this; }
  BuildConfig &enableELH() { buildMethod = ELH; return *this; }
  int maxAllowedLeafSize = 1 << 15;
  int makeLeafThreshold = 0;
  BuildMethod buildMethod = SPATIAL_MEDIAN;
};
static_assert(sizeof(BuildConfig) == sizeof(cubql_build_config), "BuildConfig layout");

template <typename T, int D> struct BinaryBVH;
template <typename T, int D, int W> struct WideBVH;

/*! BinaryBVH<float,3> (bvh.h:71-111): node 0 is the root; an inner node
    (count 0) has its children at offset and offset+1; a leaf holds
    primIDs[offset .. offset+count).  The GPU builders leave node 1 unused. */
template <> struct BinaryBVH<float, 3> {
  using scalar_t = float;
  enum { numDims = 3 };
  using vec_t = vec3f;
  using box_t = box3f;
  struct alignas(16) Node {
    enum { count_bits = 16, offset_bits = 64 - count_bits };
    box3f bounds;
    struct Admin {
      union {
        struct {
          uint64_t offset : offset_bits;
          uint64_t count : count_bits;
        };
        uint64_t offsetAndCountBits;
      };
    } admin;
  };
  enum { maxLeafSize = (1 << Node::count_bits) - 1 };
  using node_t = Node;
  Node *nodes = nullptr;
  uint32_t numNodes = 0;
  uint32_t *primIDs = nullptr;
  uint32_t numPrims = 0;
};
static_assert(sizeof(BinaryBVH<float, 3>::Node) == sizeof(cubql_node3f), "Node layout");
static_assert(sizeof(BinaryBVH<float, 3>) == sizeof(cubql_bvh3f), "BinaryBVH layout");

/*! WideBVH<float,3,W> (bvh.h:113-141): each node holds W children; a child
    is valid (bit 0), and then either an inner child (count 0, its node at
    nodes[offset]) or a leaf over primIDs[offset .. offset+count). */
template <int W> struct WideBVH<float, 3, W> {
  using scalar_t = float;
  enum { numDims = 3, width = W };
  using vec_t = vec3f;
  using box_t = box3f;
  struct alignas(16) Node {
    struct alignas(16) Child {
      box3f bounds;
      struct {
        uint64_t valid : 1;
        uint64_t offset : 45;
        uint64_t count : 16;
      };
    } children[W];
  };
  Node *nodes = nullptr;
  uint32_t numNodes = 0;
  uint32_t *primIDs = nullptr;
  uint32_t numPrims = 0;
};

template <typename T, int D> using bvh_t = BinaryBVH<T, D>;
using bvh3f = BinaryBVH<float, 3>;
using bvh_float3 = BinaryBVH<float, 3>;

/*! the C-ABI view of a bvh3f (same bytes) */
CUBQL_BOTH const cubql_bvh3f &c_bvh(const bvh3f &b) { return *reinterpret_cast<const cubql_bvh3f *>(&b); }
CUBQL_BOTH cubql_bvh3f &c_bvh(bvh3f &b) { return *reinterpret_cast<cubql_bvh3f *>(&b); }

}  // namespace cuBQL
