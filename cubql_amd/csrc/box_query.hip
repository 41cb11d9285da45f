// box_query.hip -- batched fixed-box (range) queries.
//
// Replaces user kernels that call cuBQL::fixedBoxQuery::forEachPrim once
// per thread (traversal/fixedBoxQuery.h:42-160, 215-230) with a per-prim
// test of the prim's box against the query box: for every query box, the
// prims whose boxes overlap it (closed boxes, math/box.h:208-212), in the
// traversal's order (depth first, first child / lowest slot first).  One
// query per lane; binary and wide (W = 4, 8) trees.  Queries whose stack
// would overflow rerun with a deeper one.
#include "common.h"
#include "cubql_amd/traversal.hip.h"
#include "profile.h"

namespace cubql {
namespace bq {

constexpr int kBlock = 256;
constexpr int kFastStack = 64;
constexpr int kDeepStack = 1024;

struct Overflow { uint32_t count, deepFailed; };

__device__ __forceinline__ void load_box(const cubql_box3f *b, uint32_t i, float (&lo)[3], float (&hi)[3]) {
  const float2 *f = reinterpret_cast<const float2 *>(b + i);
  const float2 x = f[0], y = f[1], z = f[2];
  lo[0] = x.x; lo[1] = x.y; lo[2] = y.x;
  hi[0] = y.y; hi[1] = z.x; hi[2] = z.y;
}

// Collects up to maxOut overlapping prims into out[]; returns false on a
// stack overflow.  *count receives the number of overlapping prims.
template <int STACK, typename BVH>
__device__ __forceinline__ bool query_one(const BVH &bvh, const cubql_box3f *primBoxes, const float (&lo)[3],
                                          const float (&hi)[3], uint32_t maxOut, uint32_t *out, uint32_t &count) {
  count = 0;
  auto leaf = [&](const uint32_t *ids, uint32_t cnt) -> int {
    for (uint32_t j = 0; j < cnt; j++) {
      const uint32_t p = ids[j];
      float bl[3], bh[3];
      load_box(primBoxes, p, bl, bh);
      const bool ov = !(bh[0] < lo[0] || bh[1] < lo[1] || bh[2] < lo[2] || hi[0] < bl[0] || hi[1] < bl[1] ||
                        hi[2] < bl[2]);
      if (ov) {
        if (count < maxOut) out[count] = p;
        count++;
      }
    }
    return 0;
  };
  return cubql_dev::for_each_leaf_in_box<STACK>(bvh, lo, hi, leaf);
}

template <typename BVH>
__global__ __launch_bounds__(kBlock) void k_box_query(BVH bvh, const cubql_box3f *primBoxes,
                                                      const cubql_box3f *queries, uint32_t nq, uint32_t maxOut,
                                                      uint32_t *outIDs, uint32_t *counts, Overflow *ovf,
                                                      uint32_t *ovfList) {
  const uint32_t qi = blockIdx.x * kBlock + threadIdx.x;
  if (qi >= nq) return;
  float lo[3], hi[3];
  load_box(queries, qi, lo, hi);
  uint32_t count;
  if (!query_one<kFastStack>(bvh, primBoxes, lo, hi, outIDs ? maxOut : 0u,
                             outIDs ? outIDs + (size_t)qi * maxOut : nullptr, count)) {
    ovfList[atomicAdd(&ovf->count, 1u)] = qi;
    return;
  }
  counts[qi] = count;
}

template <typename BVH>
__global__ __launch_bounds__(64) void k_box_query_deep(BVH bvh, const cubql_box3f *primBoxes,
                                                       const cubql_box3f *queries, uint32_t maxOut,
                                                       uint32_t *outIDs, uint32_t *counts, Overflow *ovf,
                                                       const uint32_t *ovfList) {
  const uint32_t n = ovf->count;
  for (uint32_t i = blockIdx.x * 64 + threadIdx.x; i < n; i += gridDim.x * 64) {
    const uint32_t qi = ovfList[i];
    float lo[3], hi[3];
    load_box(queries, qi, lo, hi);
    uint32_t count;
    if (!query_one<kDeepStack>(bvh, primBoxes, lo, hi, outIDs ? maxOut : 0u,
                               outIDs ? outIDs + (size_t)qi * maxOut : nullptr, count)) {
      atomicAdd(&ovf->deepFailed, 1u);
      count = 0xffffffffu;
    }
    counts[qi] = count;
  }
}

__global__ void k_zero_counts(uint32_t *counts, uint32_t nq) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x) counts[i] = 0;
}

template <typename BVH>
int run(const BVH &bvh, const cubql_box3f *primBoxes, const cubql_box3f *queries, uint32_t nq, uint32_t maxOut,
        uint32_t *outIDs, uint32_t *counts, hipStream_t s) {
  if (bvh.numNodes == 0) {
    k_zero_counts<<<std::min<uint32_t>(div_up(nq, 256), 4096), 256, 0, s>>>(counts, nq);
    CUBQL_LAUNCH_CHECK();
    return CUBQL_OK;
  }
  Scratch scr(nullptr, s);
  Overflow *ovf = nullptr;
  uint32_t *ovfList = nullptr;
  CUBQL_CHECK(scr.alloc(&ovf, 1));
  CUBQL_CHECK(scr.alloc(&ovfList, nq));
  CUBQL_CHECK(fill_u32((uint32_t *)ovf, 0u, 2, s));
  int ps = prof_begin(s);
  k_box_query<BVH><<<div_up(nq, kBlock), kBlock, 0, s>>>(bvh, primBoxes, queries, nq, maxOut, outIDs, counts, ovf,
                                                         ovfList);
  prof_end(ps, s, "box_query");
  prof_set(ps, 28.0 * nq, nq);
  k_box_query_deep<BVH><<<64, 64, 0, s>>>(bvh, primBoxes, queries, maxOut, outIDs, counts, ovf, ovfList);
  CUBQL_LAUNCH_CHECK();
  return CUBQL_OK;
}

}  // namespace bq
}  // namespace cubql

using namespace cubql;

extern "C" int cubql_box_query_f3(const cubql_bvh3f *bvh, const cubql_box3f *d_primBoxes,
                                  const cubql_box3f *d_queryBoxes, uint32_t numQueries, uint32_t maxPerQuery,
                                  uint32_t *d_outIDs, uint32_t *d_counts, cubql_stream_t stream) {
  clear_error();
  if (!bvh || (numQueries && (!d_queryBoxes || !d_counts || (bvh->numNodes && !d_primBoxes)))) {
    set_error("cubql_box_query_f3: null argument");
    return CUBQL_ERR_INVALID_ARG;
  }
  if (numQueries == 0) return CUBQL_OK;
  return bq::run(*bvh, d_primBoxes, d_queryBoxes, numQueries, maxPerQuery, d_outIDs, d_counts,
                 (hipStream_t)stream);
}

extern "C" int cubql_box_query_wide_f3(const cubql_wbvh3f *bvh, int width, const cubql_box3f *d_primBoxes,
                                       const cubql_box3f *d_queryBoxes, uint32_t numQueries, uint32_t maxPerQuery,
                                       uint32_t *d_outIDs, uint32_t *d_counts, cubql_stream_t stream) {
  clear_error();
  if (!bvh || (width != 4 && width != 8) ||
      (numQueries && (!d_queryBoxes || !d_counts || (bvh->numNodes && !d_primBoxes)))) {
    set_error("cubql_box_query_wide_f3: invalid argument (width %d)", width);
    return CUBQL_ERR_INVALID_ARG;
  }
  if (numQueries == 0) return CUBQL_OK;
  hipStream_t s = (hipStream_t)stream;
  if (width == 4)
    return bq::run(cubql_dev::WideBVH<4>{(const float4 *)bvh->nodes, bvh->numNodes, bvh->primIDs, bvh->numPrims},
                   d_primBoxes, d_queryBoxes, numQueries, maxPerQuery, d_outIDs, d_counts, s);
  return bq::run(cubql_dev::WideBVH<8>{(const float4 *)bvh->nodes, bvh->numNodes, bvh->primIDs, bvh->numPrims},
                 d_primBoxes, d_queryBoxes, numQueries, maxPerQuery, d_outIDs, d_counts, s);
}
