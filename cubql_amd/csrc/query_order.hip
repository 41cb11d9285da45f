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

// Bits of the Morton key the queries are sorted by: the top 24 of the 30
// (3 sort passes, cells of 1/256 of the root box per axis; 10M queries
// leave less than one per cell, so the last 6 bits change little).
constexpr int kOrderBits = 24;

// Top kOrderBits of the 30-bit Morton key of each query in the cells of the
// root box (queries outside it clamp to the border cells); value = query
// index.
__global__ void k_query_keys(const cubql_box3f *rootBox, const float *queries, uint32_t nq, uint32_t *keys,
                             uint32_t *vals, uint32_t *count) {
  const cubql_box3f bounds = *rootBox;
  const float ext = fmaxf(fmaxf(bounds.upper[0] - bounds.lower[0], bounds.upper[1] - bounds.lower[1]),
                          bounds.upper[2] - bounds.lower[2]);
  const float scale = 1024.f / fmaxf(ext, 1e-20f);
  if (blockIdx.x == 0 && threadIdx.x == 0) *count = nq;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x) {
    uint32_t cell[3];
    for (int d = 0; d < 3; d++) {
      const float f = fmaxf((queries[3 * (size_t)i + d] - bounds.lower[d]) * scale, 0.f);
      cell[d] = f < 1023.f ? (uint32_t)f : 1023u;
    }
    keys[i] = (spread10(cell[0]) | (spread10(cell[1]) << 1) | (spread10(cell[2]) << 2)) >> (30 - kOrderBits);
    vals[i] = i;
  }
}

}  // namespace qo

size_t query_order_ws_bytes(uint32_t nq) {
  const size_t arr = ((size_t)nq * 4 + 255) & ~(size_t)255;
  return 4 * arr + 256 + radix_sort_ws_bytes(nq, qo::kOrderBits);
}

int query_order_ws(const cubql_bvh3f &bvh, const float *queries, uint32_t nq, hipStream_t s, void *ws,
                   const uint32_t **order) {
  // the root node's box is its first 24 bytes
  return query_order_box_ws(bvh.numNodes ? &bvh.nodes[0].bounds : nullptr, queries, nq, s, ws, order);
}

int query_order_box_ws(const cubql_box3f *rootBox, const float *queries, uint32_t nq, hipStream_t s, void *ws,
                       const uint32_t **order) {
  *order = nullptr;
  if (nq < kQueryOrderMin || !rootBox) return CUBQL_OK;
  const size_t arr = ((size_t)nq * 4 + 255) & ~(size_t)255;
  char *p = (char *)ws;
  uint32_t *kA = (uint32_t *)p, *vA = (uint32_t *)(p + arr), *kB = (uint32_t *)(p + 2 * arr),
           *vB = (uint32_t *)(p + 3 * arr), *cnt = (uint32_t *)(p + 4 * arr);
  void *sortWs = p + 4 * arr + 256;
  qo::k_query_keys<<<std::min<uint32_t>(div_up(nq, 256), 4096), 256, 0, s>>>(rootBox, queries, nq, kA, vA, cnt);
  CUBQL_LAUNCH_CHECK();
  uint32_t *kRes = nullptr, *vRes = nullptr;
  CUBQL_CHECK(radix_sort_pairs_ws_u32(kA, vA, kB, vB, cnt, nq, qo::kOrderBits, s, sortWs, &kRes, &vRes));
  *order = vRes;
  return CUBQL_OK;
}

int query_order(const cubql_bvh3f &bvh, const float *queries, uint32_t nq, hipStream_t s, Scratch &scr,
                const uint32_t **order) {
  *order = nullptr;
  if (nq < kQueryOrderMin || bvh.numNodes == 0) return CUBQL_OK;
  char *ws = nullptr;
  CUBQL_CHECK(scr.alloc(&ws, query_order_ws_bytes(nq)));
  return query_order_ws(bvh, queries, nq, s, ws, order);
}

}  // namespace cubql
