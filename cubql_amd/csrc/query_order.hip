// query_order.hip -- Morton ordering of query batches (see query_order.h).
#include "query_order.h"

#include "radix_sort.h"

namespace cubql {
namespace qo {

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  uint32_t x = v & 0x3ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

// 30-bit Morton key of each query in the cells of the root box (queries
// outside it clamp to the border cells); value = query index.
__global__ void k_query_keys(const cubql_node3f *nodes, const float *queries, uint32_t nq, uint32_t *keys,
                             uint32_t *vals, uint32_t *count) {
  const cubql_node3f root = nodes[0];
  const float ext = fmaxf(fmaxf(root.bounds.upper[0] - root.bounds.lower[0],
                                root.bounds.upper[1] - root.bounds.lower[1]),
                          root.bounds.upper[2] - root.bounds.lower[2]);
  const float scale = 1024.f / fmaxf(ext, 1e-20f);
  if (blockIdx.x == 0 && threadIdx.x == 0) *count = nq;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x) {
    uint32_t cell[3];
    for (int d = 0; d < 3; d++) {
      const float f = fmaxf((queries[3 * (size_t)i + d] - root.bounds.lower[d]) * scale, 0.f);
      cell[d] = f < 1023.f ? (uint32_t)f : 1023u;
    }
    keys[i] = spread10(cell[0]) | (spread10(cell[1]) << 1) | (spread10(cell[2]) << 2);
    vals[i] = i;
  }
}

}  // namespace qo

int query_order(const cubql_bvh3f &bvh, const float *queries, uint32_t nq, hipStream_t s, Scratch &scr,
                const uint32_t **order) {
  *order = nullptr;
  if (nq < kQueryOrderMin || bvh.numNodes == 0) return CUBQL_OK;
  uint32_t *kA, *vA, *kB, *vB, *cnt;
  CUBQL_CHECK(scr.alloc(&kA, nq));
  CUBQL_CHECK(scr.alloc(&vA, nq));
  CUBQL_CHECK(scr.alloc(&kB, nq));
  CUBQL_CHECK(scr.alloc(&vB, nq));
  CUBQL_CHECK(scr.alloc(&cnt, 1));
  qo::k_query_keys<<<std::min<uint32_t>(div_up(nq, 256), 4096), 256, 0, s>>>(bvh.nodes, queries, nq, kA, vA, cnt);
  CUBQL_LAUNCH_CHECK();
  uint32_t *kRes = nullptr, *vRes = nullptr;
  CUBQL_CHECK(radix_sort_pairs_dev_u32(kA, vA, kB, vB, cnt, nq, 30, s, scr.mem, &kRes, &vRes));
  *order = vRes;
  return CUBQL_OK;
}

}  // namespace cubql
