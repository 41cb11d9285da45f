// wide_builder.hip -- collapse of a BinaryBVH into a WideBVH<float,3,W>
// (W = 4 or 8), the counterpart of gpuBuilder_impl::collapseInit /
// collapseSummarize / collapseExecute (builder/cuda/wide_gpu_builder.h:
// 39-160) and of gpuBuilder(WideBVH&, ...) (:162-215).
//
// The treelet rule is the reference's: the binary root and every inner
// binary node whose depth is a multiple of log2(W) is the root of one wide
// node; a wide node's children are the leaves and the treelet roots met by
// a depth-first walk from its root (second child first), so a wide node
// has at most W children and unused slots are invalid.  Wide node ids are
// assigned in increasing order of the binary root's index (an exclusive
// scan over the root flags), so the wide tree is a deterministic function
// of the binary tree; the reference numbers them with an atomic counter.
//
// Passes: parent links; depth by walking up and the root flag; the scan;
// one thread per wide node writing its W children.  One host
// synchronisation reads the wide node count to size the output.
#include "block_ops.h"
#include "build_common.h"
#include "common.h"

namespace cubql {
namespace wd {

constexpr uint32_t kNone = 0xffffffffu;

// GPU builders leave node 1 unused; a tree whose root has its children at 1
// and 2 (the CPU builder's numbering) uses it.
__device__ inline bool node1_used(const Node *nodes) {
  const uint64_t oc = nodes[0].oc;
  return oc_count(oc) == 0 && oc_offset(oc) == 1;
}

__global__ void k_parent(const Node *nodes, uint32_t n, uint32_t *parent) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (i == 1 && !node1_used(nodes)) continue;
    const uint64_t oc = nodes[i].oc;
    if (oc_count(oc) != 0) continue;
    const uint64_t off = oc_offset(oc);
    if (off + 1 < n) {
      parent[off] = i;
      parent[off + 1] = i;
    }
  }
}

// flag[i] = 1 iff binary node i roots a wide node
__global__ void k_flag(const Node *nodes, uint32_t n, const uint32_t *parent, int logW, uint32_t *flag) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint32_t f = 0;
    if (i == 0) {
      f = 1;
    } else if (!(i == 1 && !node1_used(nodes)) && oc_count(nodes[i].oc) == 0) {
      uint32_t depth = 0, x = i;
      // walk up to the root; a broken chain (unreachable node) or one
      // longer than the node count flags nothing
      while (x != 0 && x != kNone && depth <= n) {
        x = parent[x];
        depth++;
      }
      f = (x == 0 && depth % (uint32_t)logW == 0) ? 1u : 0u;
    }
    flag[i] = f;
  }
}

__global__ __launch_bounds__(1024) void k_scan(const uint32_t *flag, uint32_t *wideId, uint32_t n,
                                               uint64_t *status, uint32_t *total) {
  __shared__ uint32_t smem[1024 / kWave + 2];
  count_scan_tile(flag, wideId, n, 0u, status, 1u, total, smem, 0u, total + 1);
}

template <int W>
__global__ void k_exec(const Node *nodes, uint32_t n, const uint32_t *flag, const uint32_t *wideId,
                       cubql_wide_child3f *wide) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (!flag[i]) continue;
    cubql_wide_child3f *target = wide + (size_t)wideId[i] * W;
    uint32_t stack[16];
    int sp = 0, written = 0;
    stack[sp++] = i;
    while (sp > 0) {
      const uint32_t x = stack[--sp];
      const Node nd = nodes[x];
      const uint32_t cnt = oc_count(nd.oc);
      if (cnt != 0 || (x != i && flag[x])) {
        if (written < W) {
          cubql_wide_child3f c;
          for (int d = 0; d < 3; d++) {
            c.bounds.lower[d] = nd.lo[d];
            c.bounds.upper[d] = nd.hi[d];
          }
          const uint64_t off = cnt ? oc_offset(nd.oc) : (uint64_t)wideId[x];
          c.validOffsetCount = 1ull | ((off & ((1ull << 45) - 1)) << 1) | ((uint64_t)cnt << 46);
          target[written] = c;
        }
        written++;
      } else if (sp + 2 <= 16) {
        const uint32_t off = (uint32_t)oc_offset(nd.oc);
        stack[sp++] = off;
        stack[sp++] = off + 1;
      }
    }
    for (int j = written; j < W; j++) {
      cubql_wide_child3f c;
      for (int d = 0; d < 3; d++) {
        c.bounds.lower[d] = INFINITY;
        c.bounds.upper[d] = -INFINITY;
      }
      c.validOffsetCount = ((uint64_t)0xffffffffu << 1) | (0xffffull << 46);
      target[j] = c;
    }
  }
}

}  // namespace wd

// Collapse `bin` into `out` (nodes allocated from `mem`; primIDs left to the
// caller).  Blocks the host once to read the wide node count.
int collapse_wide(const cubql_bvh3f &bin, int width, cubql_wbvh3f *out, hipStream_t s,
                  const cubql_mem_resource *mem) {
  const uint32_t n = bin.numNodes;
  const int logW = width == 4 ? 2 : 3;
  out->nodes = nullptr;
  out->numNodes = 0;
  if (n == 0) return CUBQL_OK;
  ScratchArena arena(mem, s);
  uint32_t *parent, *flag, *wideId, *total;
  uint64_t *status;
  const uint32_t scanTiles = div_up(n, kCountScanTile);
  CUBQL_CHECK(arena.reserve(&parent, n));
  CUBQL_CHECK(arena.reserve(&flag, n));
  CUBQL_CHECK(arena.reserve(&wideId, n));
  CUBQL_CHECK(arena.reserve(&status, scanTiles));
  CUBQL_CHECK(arena.reserve(&total, 2));  // wide node count, look-back fault word
  CUBQL_CHECK(arena.commit());
  CUBQL_CHECK(fill_u32(parent, wd::kNone, n, s));
  CUBQL_CHECK(fill_u32(total, 0u, 2, s));
  CUBQL_CHECK(fill_u32((uint32_t *)status, 0u, 2 * (size_t)scanTiles, s));
  const uint32_t g = std::min<uint32_t>(div_up(n, 256), 16384);
  const Node *nodes = (const Node *)bin.nodes;
  wd::k_parent<<<g, 256, 0, s>>>(nodes, n, parent);
  wd::k_flag<<<g, 256, 0, s>>>(nodes, n, parent, logW, flag);
  wd::k_scan<<<scanTiles, 1024, 0, s>>>(flag, wideId, n, status, total);
  CUBQL_LAUNCH_CHECK();
  uint32_t tw[2] = {0, 0};
  CUBQL_HIP(hipMemcpyAsync(tw, total, 8, hipMemcpyDeviceToHost, s));
  CUBQL_HIP(hipStreamSynchronize(s));
  if (tw[1]) {
    set_error("wide collapse: a look-back wait exceeded its bound");
    return CUBQL_ERR_INTERNAL;
  }
  const uint32_t numWide = tw[0];
  void *wn = nullptr;
  CUBQL_CHECK(mem_alloc(mem, &wn, (size_t)numWide * width * sizeof(cubql_wide_child3f), s));
  out->nodes = (cubql_wide_child3f *)wn;
  out->numNodes = numWide;
  if (width == 4)
    wd::k_exec<4><<<g, 256, 0, s>>>(nodes, n, flag, wideId, out->nodes);
  else
    wd::k_exec<8><<<g, 256, 0, s>>>(nodes, n, flag, wideId, out->nodes);
  CUBQL_LAUNCH_CHECK();
  return CUBQL_OK;
}

}  // namespace cubql

using namespace cubql;

extern "C" int cubql_collapse_wide_f3(const cubql_bvh3f *binary, int width, cubql_wbvh3f *wide,
                                      cubql_stream_t stream, const cubql_mem_resource *mem) {
  clear_error();
  if (!binary || !wide || (width != 4 && width != 8)) {
    set_error("cubql_collapse_wide_f3: invalid argument (width %d; 4 or 8)", width);
    return CUBQL_ERR_INVALID_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  *wide = cubql_wbvh3f{nullptr, 0, nullptr, 0};
  if (binary->numNodes == 0) return CUBQL_OK;
  CUBQL_CHECK(collapse_wide(*binary, width, wide, s, mem));
  void *ids = nullptr;
  const int st = mem_alloc(mem, &ids, (size_t)binary->numPrims * 4, s);
  if (st != CUBQL_OK) {
    mem_free(mem, wide->nodes, s);
    wide->nodes = nullptr;
    wide->numNodes = 0;
    return st;
  }
  CUBQL_HIP(hipMemcpyAsync(ids, binary->primIDs, (size_t)binary->numPrims * 4, hipMemcpyDeviceToDevice, s));
  wide->primIDs = (uint32_t *)ids;
  wide->numPrims = binary->numPrims;
  return CUBQL_OK;
}

extern "C" int cubql_gpu_builder_wide_f3(cubql_wbvh3f *wide, int width, const cubql_box3f *d_boxes,
                                         uint32_t numBoxes, const cubql_build_config *cfg,
                                         cubql_stream_t stream, const cubql_mem_resource *mem) {
  clear_error();
  if (!wide || (width != 4 && width != 8)) {
    set_error("cubql_gpu_builder_wide_f3: invalid argument (width %d; 4 or 8)", width);
    return CUBQL_ERR_INVALID_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  cubql_bvh3f bin{nullptr, 0, nullptr, 0};
  CUBQL_CHECK(cubql_gpu_builder_f3(&bin, d_boxes, numBoxes, cfg, stream, mem));
  if (bin.numNodes == 0) return CUBQL_OK;
  const int st = collapse_wide(bin, width, wide, s, mem);
  if (st != CUBQL_OK) {
    cubql_free_f3(&bin, stream, mem);
    return st;
  }
  // the wide tree takes over the binary tree's primIDs (wide_gpu_builder.h:196-198)
  wide->primIDs = bin.primIDs;
  wide->numPrims = bin.numPrims;
  CUBQL_CHECK(mem_free(mem, bin.nodes, s));
  return CUBQL_OK;
}

extern "C" int cubql_free_wide_f3(cubql_wbvh3f *wide, cubql_stream_t stream, const cubql_mem_resource *mem) {
  clear_error();
  if (!wide) {
    set_error("cubql_free_wide_f3: bvh is NULL");
    return CUBQL_ERR_INVALID_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  CUBQL_CHECK(mem_free(mem, wide->primIDs, s));
  CUBQL_CHECK(mem_free(mem, wide->nodes, s));
  CUBQL_HIP(hipStreamSynchronize(s));
  *wide = cubql_wbvh3f{nullptr, 0, nullptr, 0};
  return CUBQL_OK;
}
