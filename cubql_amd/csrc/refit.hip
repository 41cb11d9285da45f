// refit.hip -- bottom-up refit of node boxes for any BinaryBVH topology.
//
// Replaces gpuBuilder_impl::refit (builder/cuda/sm_builder.h:616-693): leaf
// box = union of its prim boxes, inner box = union of its two children.
// One thread per node; leaves start, and each parent is finished by the
// second child to arrive (device-scope atomic counter per node), so every
// node is written exactly once.
#include "common.h"
#include "refit.h"

namespace cubql {
namespace rf {

// GPU builders leave node 1 unused (bvh.h:66-70); a tree whose root has its
// children at 1 and 2 (the CPU builder's numbering) uses it.
__device__ inline bool node1_used(const Node *nodes) {
  uint64_t oc = nodes[0].oc;
  return oc_count(oc) == 0 && oc_offset(oc) == 1;
}

__global__ void k_parents(const Node *nodes, uint32_t numNodes, uint32_t *parent, uint32_t *arrive) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < numNodes; i += gridDim.x * blockDim.x) {
    arrive[i] = 0;
    if (i == 0) parent[0] = 0xffffffffu;
    if (i == 1 && !node1_used(nodes)) continue;
    uint64_t oc = nodes[i].oc;
    if (oc_count(oc) != 0) continue;
    uint32_t off = (uint32_t)oc_offset(oc);
    if (off + 1 < numNodes) {
      parent[off] = i;
      parent[off + 1] = i;
    }
  }
}

__device__ inline void load_box(const cubql_box3f *p, uint32_t i, float lo[3], float hi[3]) {
  const float2 *f = reinterpret_cast<const float2 *>(p + i);
  float2 a = f[0], b = f[1], c = f[2];
  lo[0] = a.x; lo[1] = a.y; lo[2] = b.x;
  hi[0] = b.y; hi[1] = c.x; hi[2] = c.y;
}

__global__ void k_refit(Node *nodes, uint32_t numNodes, const uint32_t *primIDs,
                        const cubql_box3f *boxes, const uint32_t *parent, uint32_t *arrive) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < numNodes; i += gridDim.x * blockDim.x) {
    if (i == 1 && !node1_used(nodes)) continue;
    uint64_t oc = nodes[i].oc;
    uint32_t cnt = oc_count(oc);
    if (cnt == 0) continue;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    uint32_t off = (uint32_t)oc_offset(oc);
    for (uint32_t j = 0; j < cnt; j++) {
      float bl[3], bh[3];
      load_box(boxes, primIDs[off + j], bl, bh);
      for (int d = 0; d < 3; d++) {
        lo[d] = fminf(lo[d], bl[d]);
        hi[d] = fmaxf(hi[d], bh[d]);
      }
    }
    uint32_t node = i;
    while (true) {
      Node *nd = &nodes[node];
      for (int d = 0; d < 3; d++) {
        nd->lo[d] = lo[d];
        nd->hi[d] = hi[d];
      }
      uint32_t p = parent[node];
      if (p == 0xffffffffu) break;
      __threadfence();  // release this node's box before signalling
      if (atomicAdd(&arrive[p], 1u) == 0) break;  // sibling finishes the parent
      __threadfence();  // acquire the sibling's box
      uint32_t off2 = (uint32_t)oc_offset(nodes[p].oc);
      const volatile Node *l = &nodes[off2];
      const volatile Node *r = &nodes[off2 + 1];
      for (int d = 0; d < 3; d++) {
        lo[d] = fminf(l->lo[d], r->lo[d]);
        hi[d] = fmaxf(l->hi[d], r->hi[d]);
      }
      node = p;
    }
  }
}

}  // namespace rf

int refit_tree(cubql_bvh3f *bvh, const cubql_box3f *boxes, hipStream_t s,
               const cubql_mem_resource *mem) {
  if (bvh->numNodes == 0) return CUBQL_OK;
  Scratch scr(mem, s);
  uint32_t *parent = nullptr, *arrive = nullptr;
  CUBQL_CHECK(scr.alloc(&parent, bvh->numNodes));
  CUBQL_CHECK(scr.alloc(&arrive, bvh->numNodes));
  uint32_t g = std::min<uint32_t>(div_up(bvh->numNodes, 256), 8192);
  rf::k_parents<<<g, 256, 0, s>>>((const Node *)bvh->nodes, bvh->numNodes, parent, arrive);
  rf::k_refit<<<g, 256, 0, s>>>((Node *)bvh->nodes, bvh->numNodes, bvh->primIDs, boxes, parent, arrive);
  CUBQL_LAUNCH_CHECK();
  return CUBQL_OK;
}

}  // namespace cubql

extern "C" int cubql_refit_f3(cubql_bvh3f *bvh, const cubql_box3f *d_boxes, cubql_stream_t s,
                              const cubql_mem_resource *mem) {
  cubql::clear_error();
  if (!bvh || !d_boxes) {
    cubql::set_error("cubql_refit_f3: null argument");
    return CUBQL_ERR_INVALID_ARG;
  }
  return cubql::refit_tree(bvh, d_boxes, (hipStream_t)s, mem);
}
