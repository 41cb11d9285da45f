// points_query.hip -- batched closest-point and kNN queries over a point BVH.
//
// Replaces user kernels that call cuBQL::points::findClosest / findKNN once
// per thread (queries/points/findClosest.h:181-226, queries/points/knn.h:
// 75-116; sample launch: samples/s01_closestPoint_points_gpu/
// closestPoint.cu:41-66).  One query per lane of a wave64; traversal from
// include/cubql_amd/traversal.hip.h.  Result rule: among equal float32
// distances the lowest primID wins (kNN lists are ordered by
// (sqrDist, primID)).
#include <mutex>
#include <type_traits>

#include "common.h"
#include "cubql_amd/traversal.hip.h"
#include "profile.h"
#include "query_order.h"

namespace cubql {
namespace pq {

constexpr int kBlock = 256;
constexpr int kFastStack = 64;
constexpr int kDeepStack = 1024;
// id of a query whose traversal needed more than kDeepStack entries
constexpr int32_t kIncomplete = -2;

__device__ __forceinline__ float3 load_pt(const float *p, uint32_t i) {
  return make_float3(p[3 * (size_t)i + 0], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2]);
}

struct Overflow {
  uint32_t count;
  uint32_t deepFailed;
};

template <int STACK, typename BVH, typename Stats = cubql_dev::NoStats>
__device__ __forceinline__ bool fcp_one(const BVH &bvh, const float *points, float3 q,
                                        float sqrMax, int32_t &best, float &bestD, Stats &&st = Stats()) {
  best = -1;
  bestD = sqrMax;
  auto prim = [&](uint32_t pid) -> float {
    st.prim();
    float3 p = load_pt(points, pid);
    float d = cubql_dev::sqr_dist_pt(p.x, p.y, p.z, q.x, q.y, q.z);
    if (d < bestD || (d == bestD && best >= 0 && (int32_t)pid < best)) {
      bestD = d;
      best = (int32_t)pid;
    }
    return bestD;
  };
  // seed the radius with the nearest-child leaf, which the traversal then
  // leaves out
  const uint32_t seed = cubql_dev::visit_nearest_leaf(bvh, q.x, q.y, q.z, prim, 0u, st);
  return cubql_dev::for_each_prim_except<STACK>(bvh, q.x, q.y, q.z, bestD, prim, seed, 0u, st);
}

// The same over a binary tree on a caller-provided traversal stack.
template <typename StackT>
__device__ __forceinline__ bool fcp_one_on(StackT &stack, const cubql_bvh3f &bvh, const float *points, float3 q,
                                           float sqrMax, int32_t &best, float &bestD) {
  best = -1;
  bestD = sqrMax;
  auto prim = [&](uint32_t pid) -> float {
    float3 p = load_pt(points, pid);
    float d = cubql_dev::sqr_dist_pt(p.x, p.y, p.z, q.x, q.y, q.z);
    if (d < bestD || (d == bestD && best >= 0 && (int32_t)pid < best)) {
      bestD = d;
      best = (int32_t)pid;
    }
    return bestD;
  };
  cubql_dev::NoStats st;
  const uint32_t seed = cubql_dev::visit_nearest_leaf(bvh, q.x, q.y, q.z, prim, 0u, st);
  return cubql_dev::for_each_prim_except_on(stack, bvh, q.x, q.y, q.z, bestD, prim, seed, 0u, st);
}

// The same with the leaf prims read from a copy of the points in primID-slot
// order (x, y, z and the primID's bits): a leaf's prims are one contiguous
// run instead of a primID load and a gather each.  Tests the same prims in
// the same order, so the result is the same.
template <typename StackT>
__device__ __forceinline__ bool fcp_sorted_on(StackT &stack, const cubql_bvh3f &bvh, const float4 *spts, float3 q,
                                              float sqrMax, int32_t &best, float &bestD) {
  best = -1;
  bestD = sqrMax;
  auto test = [&](uint32_t slot) {
    const float4 p = spts[slot];
    const float d = cubql_dev::sqr_dist_pt(p.x, p.y, p.z, q.x, q.y, q.z);
    const int32_t pid = (int32_t)__float_as_uint(p.w);
    if (d < bestD || (d == bestD && best >= 0 && pid < best)) {
      bestD = d;
      best = pid;
    }
  };
  cubql_dev::NoStats st;
  const uint64_t seedLeaf = cubql_dev::nearest_leaf(bvh, q.x, q.y, q.z, 0u, st);
  const uint32_t seed = cubql_dev::oc_off(seedLeaf);
  for (uint32_t j = 0; j < cubql_dev::oc_cnt(seedLeaf); j++) test(seed + j);
  auto leaf = [&](const uint32_t *ids, uint32_t cnt) -> float {
    const uint32_t off = (uint32_t)(ids - bvh.primIDs);
    if (off == seed) return INFINITY;
    for (uint32_t j = 0; j < cnt; j++) test(off + j);
    return bestD;
  };
  return cubql_dev::for_each_leaf_on(stack, bvh, q.x, q.y, q.z, bestD, leaf, 0u, st);
}

// points in primID-slot order for fcp_sorted_on
__global__ __launch_bounds__(256) void k_slot_points(const uint32_t *primIDs, uint32_t numPrims, const float *points,
                                                     float4 *spts) {
  for (uint32_t t = blockIdx.x * 256 + threadIdx.x; t < numPrims; t += gridDim.x * 256) {
    const uint32_t i = primIDs[t];
    const float3 p = load_pt(points, i);
    spts[t] = make_float4(p.x, p.y, p.z, __uint_as_float(i));
  }
}

// Entries of the fast kernel's LDS stack per lane: the seeded radius keeps
// a closest-point query's stack shallow, and 8 entries x 8 bytes x 256 lanes
// (16 KB a workgroup) leave the kernel at full occupancy; deeper queries
// rerun in the deep kernel.
constexpr int kLdsStack = 8;

// fast path: an 8-entry LDS stack over a binary tree, a 64-entry private
// stack over a wide one; queries that overflow go to the deep kernel.
// BVH: cubql_bvh3f (binary) or cubql_dev::WideBVH<W>.
// kSorted (binary trees): leaf prims from spts, the points in primID-slot order.
template <typename BVH, bool kSorted = false>
__global__ __launch_bounds__(kBlock) void k_fcp(BVH bvh, const float *points, const float4 *spts,
                                                const float *queries, uint32_t nq, float sqrMax,
                                                int32_t *outIds, float *outD, Overflow *ovf,
                                                uint32_t *ovfList, const uint32_t *order) {
  // Morton-ordered queries: one contiguous run of them per XCD
  const uint32_t i = (order ? xcd_chunk(blockIdx.x, gridDim.x) : blockIdx.x) * kBlock + threadIdx.x;
  if (i >= nq) return;
  const uint32_t qi = order ? order[i] : i;
  float3 q = load_pt(queries, qi);
  int32_t best;
  float bestD;
  bool done;
  if constexpr (std::is_same<BVH, cubql_bvh3f>::value) {
    __shared__ uint2 sStack[kLdsStack * kBlock];
    cubql_dev::LdsStack<kLdsStack, kBlock> stack(sStack + threadIdx.x);
    if constexpr (kSorted) done = fcp_sorted_on(stack, bvh, spts, q, sqrMax, best, bestD);
    else done = fcp_one_on(stack, bvh, points, q, sqrMax, best, bestD);
  } else {
    done = fcp_one<kFastStack>(bvh, points, q, sqrMax, best, bestD);
  }
  if (!done) {
    uint32_t slot = atomicAdd(&ovf->count, 1u);
    ovfList[slot] = qi;
    return;
  }
  outIds[qi] = best;
  if (outD) outD[qi] = best >= 0 ? bestD : INFINITY;
}

template <typename BVH>
__global__ __launch_bounds__(64) void k_fcp_deep(BVH bvh, const float *points,
                                                 const float *queries, float sqrMax,
                                                 int32_t *outIds, float *outD, Overflow *ovf,
                                                 const uint32_t *ovfList) {
  const uint32_t n = ovf->count;
  for (uint32_t i = blockIdx.x * 64 + threadIdx.x; i < n; i += gridDim.x * 64) {
    const uint32_t qi = ovfList[i];
    float3 q = load_pt(queries, qi);
    int32_t best;
    float bestD;
    if (!fcp_one<kDeepStack>(bvh, points, q, sqrMax, best, bestD)) {
      // deeper than the deep stack: marked incomplete, not "none found"
      atomicAdd(&ovf->deepFailed, 1u);
      best = kIncomplete;
    }
    outIds[qi] = best;
    if (outD) outD[qi] = best >= 0 ? bestD : INFINITY;
  }
}

// kNN with a register-resident candidate list of K entries (K >= k).
// kSorted (binary trees): leaf prims from spts, the points in primID-slot
// order, a contiguous run per leaf (the same prims in the same order).
template <int K, int STACK, bool kSorted = false, typename BVH, typename Stats = cubql_dev::NoStats>
__device__ __forceinline__ bool knn_one(const BVH &bvh, const float *points, const float4 *spts, float3 q,
                                        int k, float sqrMax, int32_t (&ids)[K], float (&ds)[K],
                                        int &found, Stats &&st = Stats()) {
  found = 0;
#pragma unroll
  for (int j = 0; j < K; j++) { ids[j] = -1; ds[j] = INFINITY; }
  float cull = sqrMax;
  // (ds[k-1], ids[k-1]) once the list is full, kept in registers so that a
  // rejected candidate (most of them) costs two compares, not a K-wide
  // select to read the runtime-indexed k-th entry
  float kthD = INFINITY;
  int32_t kthId = 0x7fffffff;
  auto insert = [&](float3 p, uint32_t pid) -> float {
    st.prim();
    float d = cubql_dev::sqr_dist_pt(p.x, p.y, p.z, q.x, q.y, q.z);
    int32_t id = (int32_t)pid;
    if (!(d < sqrMax)) return cull;
    if (found == k) {
      // reject unless (d,id) < (ds[k-1], ids[k-1])
      if (d > kthD || (d == kthD && id > kthId)) return cull;
    } else {
      found++;
    }
    // insertion: shift entries greater than (d,id) one slot to the right,
    // over the first `found` slots (register-unrolled)
    int32_t carryI = id;
    float carryD = d;
#pragma unroll
    for (int j = 0; j < K; j++) {
      if (j < found) {
        bool greater = ds[j] > carryD || (ds[j] == carryD && ids[j] > carryI);
        if (greater) {
          float td = ds[j];
          int32_t ti = ids[j];
          ds[j] = carryD;
          ids[j] = carryI;
          carryD = td;
          carryI = ti;
        }
      }
    }
    if (found == k) {
#pragma unroll
      for (int j = 0; j < K; j++)
        if (j == k - 1) {
          kthD = ds[j];
          kthId = ids[j];
        }
      cull = kthD;
    }
    return cull;
  };
  if constexpr (kSorted) {
    auto leaf = [&](const uint32_t *idp, uint32_t cnt) -> float {
      const uint32_t off = (uint32_t)(idp - bvh.primIDs);
      float r = INFINITY;
      for (uint32_t c = 0; c < cnt; c++) {
        const float4 v = spts[off + c];
        r = fminf(r, insert(make_float3(v.x, v.y, v.z), __float_as_uint(v.w)));
      }
      return r;
    };
    return cubql_dev::for_each_leaf<STACK>(bvh, q.x, q.y, q.z, sqrMax, leaf, 0u, st);
  } else {
    auto prim = [&](uint32_t pid) -> float { return insert(load_pt(points, pid), pid); };
    return cubql_dev::for_each_prim<STACK>(bvh, q.x, q.y, q.z, sqrMax, prim, 0u, st);
  }
}

// Far siblings of the nearest leaf's last kSeedSibs ancestors whose nearest
// leaves seed a kNN query over a binary tree.
constexpr int kSeedSibs = 4;

// kNN over a binary tree on a caller-provided (short) stack.  Before the
// traversal proper, the query visits the nearest leaf (descending along the
// nearer child, ties to the second) and then, while its list is not full,
// the nearest leaf under each far sibling met on the last kSeedSibs steps of
// that descent (deepest first).  The traversal then starts at the root with
// the list's radius (the k-th distance once the list is full), so that on
// the way down it stacks only far children that can still hold a candidate,
// and it skips the seeded leaves, whose prims are already in the list.  The
// result is knn_one's: the list holds the k smallest (distance, primID)
// pairs under sqrMax whatever the visiting order, and every leaf within the
// radius is still visited exactly once.  One loop visits every leaf, so the
// insertion code is inlined once.
template <int K, bool kSorted = false, typename StackT, typename Stats = cubql_dev::NoStats>
__device__ __forceinline__ bool knn_seeded_on(StackT &stack, const cubql_bvh3f &bvh, const float *points,
                                              const float4 *spts, float3 q, int k, float sqrMax,
                                              int32_t (&ids)[K], float (&ds)[K], int &found,
                                              Stats &&st = Stats()) {
  using namespace cubql_dev;
  found = 0;
#pragma unroll
  for (int j = 0; j < K; j++) { ids[j] = -1; ds[j] = INFINITY; }
  float cull = sqrMax;
  float kthD = INFINITY;
  int32_t kthId = 0x7fffffff;
  // far siblings of the first descent (sib[0] the deepest) and the offsets
  // of the seeded leaves (a leaf's first primID slot names it), the latest
  // in seeded[0]; every index is static, so both stay in registers
  uint32_t sib[kSeedSibs], seeded[kSeedSibs + 1];
#pragma unroll
  for (int j = 0; j < kSeedSibs; j++) sib[j] = 0u;
#pragma unroll
  for (int j = 0; j <= kSeedSibs; j++) seeded[j] = 0xffffffffu;
  int nsib = 0;
  bool trav = false, record = true, visit = true;
  int sp = 0;
  uint64_t node = node_oc(load_node(bvh.nodes, 0));
  while (true) {
    if (!trav) {
      // seeding descent to the nearest leaf (the first one records the far
      // siblings)
      while (oc_cnt(node) == 0) {
        const uint32_t i0 = oc_off(node);
        const QNode n0 = load_node(bvh.nodes, i0);
        const QNode n1 = load_node(bvh.nodes, i0 + 1);
        st.nodes(64);
        const bool near0 = sqr_dist_box(q.x, q.y, q.z, n0) < sqr_dist_box(q.x, q.y, q.z, n1);
        node = near0 ? node_oc(n0) : node_oc(n1);
        if (record) {
#pragma unroll
          for (int j = kSeedSibs - 1; j > 0; j--) sib[j] = sib[j - 1];
          sib[0] = near0 ? i0 + 1 : i0;
          nsib = min(nsib + 1, kSeedSibs);
        }
      }
      record = false;
    }
    if (visit) {
      const uint32_t off = oc_off(node), cnt = oc_cnt(node);
      bool skip = false;
      if (trav) {
#pragma unroll
        for (int j = 0; j <= kSeedSibs; j++) skip |= off == seeded[j];
      } else {
#pragma unroll
        for (int j = kSeedSibs; j > 0; j--) seeded[j] = seeded[j - 1];
        seeded[0] = off;
      }
      if (!skip)
        for (uint32_t c = 0; c < cnt; c++) {
          uint32_t pid;
          float3 p;
          if constexpr (kSorted) {
            const float4 v = spts[off + c];
            pid = __float_as_uint(v.w);
            p = make_float3(v.x, v.y, v.z);
          } else {
            pid = bvh.primIDs[off + c];
            p = load_pt(points, pid);
          }
          st.prim();
          const float d = sqr_dist_pt(p.x, p.y, p.z, q.x, q.y, q.z);
          const int32_t id = (int32_t)pid;
          if (!(d < sqrMax)) continue;
          if (found == k) {
            // reject unless (d,id) < (ds[k-1], ids[k-1])
            if (d > kthD || (d == kthD && id > kthId)) continue;
          } else {
            found++;
          }
          int32_t carryI = id;
          float carryD = d;
#pragma unroll
          for (int j = 0; j < K; j++) {
            if (j < found) {
              const bool greater = ds[j] > carryD || (ds[j] == carryD && ids[j] > carryI);
              if (greater) {
                const float td = ds[j];
                const int32_t ti = ids[j];
                ds[j] = carryD;
                ids[j] = carryI;
                carryD = td;
                carryI = ti;
              }
            }
          }
          if (found == k) {
#pragma unroll
            for (int j = 0; j < K; j++)
              if (j == k - 1) {
                kthD = ds[j];
                kthId = ids[j];
              }
            cull = kthD;
          }
        }
    }
    if (!trav) {
      if (nsib > 0 && found < k) {
        // next seed: the nearest leaf under the deepest unused sibling
        const uint32_t from = sib[0];
#pragma unroll
        for (int j = 0; j + 1 < kSeedSibs; j++) sib[j] = sib[j + 1];
        nsib--;
        node = node_oc(load_node(bvh.nodes, from));
        st.nodes(32);
        visit = true;
        continue;
      }
      trav = true;
      node = node_oc(load_node(bvh.nodes, 0));
    } else {
      bool popped = false;
      while (sp > 0) {
        const uint2 e = stack.get(--sp);
        if (__uint_as_float(e.y) <= cull) {
          node = node_oc(load_node(bvh.nodes, e.x));
          st.nodes(32);
          popped = true;
          break;
        }
      }
      if (!popped) return true;
    }
    // traversal descent: near child first, far child stacked if within the
    // radius (the rules of for_each_leaf_on)
    visit = true;
    while (oc_cnt(node) == 0) {
      const uint32_t i0 = oc_off(node);
      const QNode n0 = load_node(bvh.nodes, i0);
      const QNode n1 = load_node(bvh.nodes, i0 + 1);
      st.nodes(64);
      const float d0 = sqr_dist_box(q.x, q.y, q.z, n0);
      const float d1 = sqr_dist_box(q.x, q.y, q.z, n1);
      if (fminf(d0, d1) > cull) { visit = false; break; }
      const bool near0 = d0 < d1;
      node = near0 ? node_oc(n0) : node_oc(n1);
      const float dfar = fmaxf(d0, d1);
      if (dfar <= cull) {
        if (sp >= StackT::kCap) return false;
        stack.put(sp++, make_uint2(near0 ? i0 + 1 : i0, __float_as_uint(dfar)));
      }
    }
  }
}

// kNN lists of up to kKnnSeededMax entries over a binary tree take the
// seeded traversal on an LDS stack of kKnnLdsStack entries per lane (32 KB a
// workgroup; deeper queries rerun in the deep kernel).  Longer lists keep
// the unseeded traversal on a private stack: on the clustered Gaussian-
// mixture workload (2M points, k = 16) the LDS versions were slower (16
// entries: 264 Mq/s, 24 entries: 224 Mq/s, against 293 Mq/s), the queries
// that overflow the LDS stack being the slowest ones, which the deep rerun
// then no longer overlaps with the others.
constexpr int kKnnSeededMax = 8;
constexpr int kKnnLdsStack = 16;

// kSorted (binary trees): leaf prims from spts, the points in primID-slot order.
template <int K, typename BVH, bool kSorted = false>
__global__ __launch_bounds__(kBlock) void k_knn(BVH bvh, const float *points, const float4 *spts,
                                                const float *queries, uint32_t nq, int k,
                                                float sqrMax, int32_t *outIds, float *outD,
                                                int32_t *outFound, Overflow *ovf,
                                                uint32_t *ovfList, const uint32_t *order) {
  // Morton-ordered queries: one contiguous run of them per XCD
  const uint32_t i = (order ? xcd_chunk(blockIdx.x, gridDim.x) : blockIdx.x) * kBlock + threadIdx.x;
  if (i >= nq) return;
  const uint32_t qi = order ? order[i] : i;
  float3 q = load_pt(queries, qi);
  int32_t ids[K];
  float ds[K];
  int found;
  bool done;
  if constexpr (std::is_same<BVH, cubql_bvh3f>::value && K <= kKnnSeededMax) {
    __shared__ uint2 sStack[kKnnLdsStack * kBlock];
    cubql_dev::LdsStack<kKnnLdsStack, kBlock> stack(sStack + threadIdx.x);
    done = knn_seeded_on<K, kSorted>(stack, bvh, points, spts, q, k, sqrMax, ids, ds, found);
  } else {
    done = knn_one<K, kFastStack, kSorted>(bvh, points, spts, q, k, sqrMax, ids, ds, found);
  }
  if (!done) {
    uint32_t slot = atomicAdd(&ovf->count, 1u);
    ovfList[slot] = qi;
    return;
  }
#pragma unroll
  for (int j = 0; j < K; j++) {
    if (j < k) {
      outIds[(size_t)qi * k + j] = ids[j];
      outD[(size_t)qi * k + j] = ds[j];
    }
  }
  if (outFound) outFound[qi] = found;
}

template <int K, typename BVH>
__global__ __launch_bounds__(64) void k_knn_deep(BVH bvh, const float *points,
                                                 const float *queries, int k, float sqrMax,
                                                 int32_t *outIds, float *outD, int32_t *outFound,
                                                 Overflow *ovf, const uint32_t *ovfList) {
  const uint32_t n = ovf->count;
  for (uint32_t i = blockIdx.x * 64 + threadIdx.x; i < n; i += gridDim.x * 64) {
    const uint32_t qi = ovfList[i];
    float3 q = load_pt(queries, qi);
    int32_t ids[K];
    float ds[K];
    int found;
    bool done;
    if constexpr (std::is_same<BVH, cubql_bvh3f>::value && K <= kKnnSeededMax) {
      cubql_dev::PrivateStack<kDeepStack> stack;
      done = knn_seeded_on<K>(stack, bvh, points, nullptr, q, k, sqrMax, ids, ds, found);
    } else {
      done = knn_one<K, kDeepStack>(bvh, points, nullptr, q, k, sqrMax, ids, ds, found);
    }
    if (!done) {
      // deeper than the deep stack: marked incomplete, not "none found"
      atomicAdd(&ovf->deepFailed, 1u);
      found = -1;
      for (int j = 0; j < K; j++) { ids[j] = kIncomplete; ds[j] = INFINITY; }
    }
    for (int j = 0; j < K; j++) {
      if (j < k) {
        outIds[(size_t)qi * k + j] = ids[j];
        outD[(size_t)qi * k + j] = ds[j];
      }
    }
    if (outFound) outFound[qi] = found;
  }
}

__global__ void k_fill_notfound(uint32_t nq, int k, int32_t *ids, float *d, int32_t *found) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t total = (size_t)nq * k;
  for (; i < total; i += (size_t)gridDim.x * blockDim.x) {
    ids[i] = -1;
    if (d) d[i] = INFINITY;
    if (found && i < nq) found[i] = 0;
  }
}

template <int K, typename BVH>
int launch_knn(const BVH &bvh, const float *points, const float *queries, uint32_t nq,
               int k, float sqrMax, int32_t *ids, float *d, int32_t *found, Overflow *ovf,
               uint32_t *ovfList, const uint32_t *order, float4 *spts, hipStream_t s) {
  int ps = prof_begin(s);
  bool sorted = false;
  if constexpr (std::is_same<BVH, cubql_bvh3f>::value) {
    if (spts) {
      k_slot_points<<<std::max(1u, std::min(div_up(bvh.numPrims, 256u), 2048u)), 256, 0, s>>>(
          bvh.primIDs, bvh.numPrims, points, spts);
      k_knn<K, BVH, true><<<div_up(nq, kBlock), kBlock, 0, s>>>(bvh, points, spts, queries, nq, k, sqrMax, ids, d,
                                                                found, ovf, ovfList, order);
      sorted = true;
    }
  }
  if (!sorted) {
    k_knn<K, BVH><<<div_up(nq, kBlock), kBlock, 0, s>>>(bvh, points, nullptr, queries, nq, k, sqrMax, ids, d, found,
                                                        ovf, ovfList, order);
  }
  prof_end(ps, s, "knn_points");
  prof_set(ps, (12.0 + 8.0 * k + 4.0) * nq, nq);
  // one workgroup per CU: the overflow list holds a few percent of the
  // queries on clustered input
  k_knn_deep<K, BVH><<<256, 64, 0, s>>>(bvh, points, queries, k, sqrMax, ids, d, found, ovf, ovfList);
  CUBQL_LAUNCH_CHECK();
  return CUBQL_OK;
}


// union of the valid children of wide node 0: the root box the query
// ordering quantizes queries in
template <int W>
__global__ void k_wide_root_box(const cubql_wide_child3f *nodes, cubql_box3f *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  cubql_box3f b;
  for (int d = 0; d < 3; d++) {
    b.lower[d] = INFINITY;
    b.upper[d] = -INFINITY;
  }
  for (int j = 0; j < W; j++) {
    if (!(nodes[j].validOffsetCount & 1ull)) continue;
    for (int d = 0; d < 3; d++) {
      b.lower[d] = fminf(b.lower[d], nodes[j].bounds.lower[d]);
      b.upper[d] = fmaxf(b.upper[d], nodes[j].bounds.upper[d]);
    }
  }
  *out = b;
}

// Device view and root box of a tree: the binary root's box is the first 24
// bytes of node 0; a wide tree's is reduced into scratch.
inline int root_box(const cubql_bvh3f &bvh, hipStream_t, Scratch &, const cubql_box3f **box) {
  *box = &bvh.nodes[0].bounds;
  return CUBQL_OK;
}
template <int W>
int root_box(const cubql_dev::WideBVH<W> &bvh, hipStream_t s, Scratch &scr, const cubql_box3f **box) {
  cubql_box3f *b = nullptr;
  CUBQL_CHECK(scr.alloc(&b, 1));
  k_wide_root_box<W><<<1, 64, 0, s>>>((const cubql_wide_child3f *)bvh.nodes, b);
  CUBQL_LAUNCH_CHECK();
  *box = b;
  return CUBQL_OK;
}

// Workspace of one closest-point call over nq queries (256-byte aligned
// slots): overflow counters | root box | overflow list | query order.
constexpr size_t kWsOvf = 0, kWsRoot = 128, kWsList = 256;
inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
inline size_t fcp_ws_bytes(uint32_t nq) {
  return kWsList + align256((size_t)nq * 4) + (nq >= kQueryOrderMin ? query_order_ws_bytes(nq) : 0);
}

inline const cubql_box3f *root_box_ws(const cubql_bvh3f &bvh, hipStream_t, char *) { return &bvh.nodes[0].bounds; }
template <int W>
const cubql_box3f *root_box_ws(const cubql_dev::WideBVH<W> &bvh, hipStream_t s, char *ws) {
  cubql_box3f *b = (cubql_box3f *)(ws + kWsRoot);
  k_wide_root_box<W><<<1, 64, 0, s>>>((const cubql_wide_child3f *)bvh.nodes, b);
  return b;
}

// Closest-point queries on a caller-provided workspace of fcp_ws_bytes(nq)
// bytes; allocates nothing.  With spts (binary trees: numPrims float4 slots)
// the points are first copied into primID-slot order and the leaves read
// them from there.
template <typename BVH>
int fcp_run_ws(const BVH &bvh, const float *d_points, const float *d_queries, uint32_t numQueries,
               float sqrMaxSearchDist, int32_t *d_outIDs, float *d_outSqrDist, hipStream_t s, char *ws,
               const char *profName, float4 *spts = nullptr) {
  if (bvh.numNodes == 0) {
    k_fill_notfound<<<256, 256, 0, s>>>(numQueries, 1, d_outIDs, d_outSqrDist, nullptr);
    CUBQL_LAUNCH_CHECK();
    return CUBQL_OK;
  }
  Overflow *ovf = (Overflow *)(ws + kWsOvf);
  uint32_t *ovfList = (uint32_t *)(ws + kWsList);
  // a fill kernel, not a memset node: callers may capture this in a graph
  CUBQL_CHECK(fill_u32((uint32_t *)ovf, 0u, 2, s));
  const cubql_box3f *rootBox = root_box_ws(bvh, s, ws);
  CUBQL_LAUNCH_CHECK();
  const uint32_t *order = nullptr;
  if (numQueries >= kQueryOrderMin)
    CUBQL_CHECK(query_order_box_ws(rootBox, d_queries, numQueries, s,
                                   ws + kWsList + align256((size_t)numQueries * 4), &order));
  int ps = prof_begin(s);
  bool sorted = false;
  if constexpr (std::is_same<BVH, cubql_bvh3f>::value) {
    if (spts) {
      k_slot_points<<<std::max(1u, std::min(div_up(bvh.numPrims, 256u), 2048u)), 256, 0, s>>>(
          bvh.primIDs, bvh.numPrims, d_points, spts);
      k_fcp<BVH, true><<<div_up(numQueries, kBlock), kBlock, 0, s>>>(bvh, d_points, spts, d_queries, numQueries,
                                                                     sqrMaxSearchDist, d_outIDs, d_outSqrDist, ovf,
                                                                     ovfList, order);
      sorted = true;
    }
  }
  if (!sorted) {
    k_fcp<BVH><<<div_up(numQueries, kBlock), kBlock, 0, s>>>(bvh, d_points, nullptr, d_queries, numQueries,
                                                             sqrMaxSearchDist, d_outIDs, d_outSqrDist, ovf, ovfList,
                                                             order);
  }
  prof_end(ps, s, profName);
  // compulsory bytes only: query in, id + distance out
  prof_set(ps, 20.0 * numQueries, numQueries);
  k_fcp_deep<BVH><<<64, 64, 0, s>>>(bvh, d_points, d_queries, sqrMaxSearchDist, d_outIDs, d_outSqrDist, ovf,
                                     ovfList);
  CUBQL_LAUNCH_CHECK();
  return CUBQL_OK;
}

template <typename BVH>
int fcp_run(const BVH &bvh, const float *d_points, const float *d_queries, uint32_t numQueries,
            float sqrMaxSearchDist, int32_t *d_outIDs, float *d_outSqrDist, hipStream_t s, const char *profName) {
  Scratch scr(nullptr, s);
  char *ws = nullptr;
  if (bvh.numNodes) CUBQL_CHECK(scr.alloc(&ws, fcp_ws_bytes(numQueries)));
  // the slot-ordered copy pays when the queries test at least about as many
  // prims as the copy moves
  float4 *spts = nullptr;
  if constexpr (std::is_same<BVH, cubql_bvh3f>::value) {
    if (bvh.numNodes && bvh.numPrims && numQueries >= kQueryOrderMin && (uint64_t)numQueries * 8 >= bvh.numPrims)
      CUBQL_CHECK(scr.alloc(&spts, bvh.numPrims));
  }
  return fcp_run_ws(bvh, d_points, d_queries, numQueries, sqrMaxSearchDist, d_outIDs, d_outSqrDist, s, ws,
                    profName, spts);
}

template <typename BVH>
int knn_run(const BVH &bvh, const float *d_points, const float *d_queries, uint32_t numQueries, int k,
            float sqrMaxSearchDist, int32_t *d_outIDs, float *d_outSqrDist, int32_t *d_numFound, hipStream_t s) {
  if (bvh.numNodes == 0) {
    k_fill_notfound<<<256, 256, 0, s>>>(numQueries, k, d_outIDs, d_outSqrDist, d_numFound);
    CUBQL_LAUNCH_CHECK();
    return CUBQL_OK;
  }
  Scratch scr(nullptr, s);
  Overflow *ovf = nullptr;
  uint32_t *ovfList = nullptr;
  CUBQL_CHECK(scr.alloc(&ovf, 1));
  CUBQL_CHECK(scr.alloc(&ovfList, numQueries));
  CUBQL_CHECK(fill_u32((uint32_t *)ovf, 0u, 2, s));
  const cubql_box3f *rootBox = nullptr;
  CUBQL_CHECK(root_box(bvh, s, scr, &rootBox));
  const uint32_t *order = nullptr;
  if (numQueries >= kQueryOrderMin) {
    char *w = nullptr;
    CUBQL_CHECK(scr.alloc(&w, query_order_ws_bytes(numQueries)));
    CUBQL_CHECK(query_order_box_ws(rootBox, d_queries, numQueries, s, w, &order));
  }
  // the slot-ordered copy pays when the queries test at least about as many
  // prims as the copy moves
  float4 *spts = nullptr;
  if constexpr (std::is_same<BVH, cubql_bvh3f>::value) {
    if (bvh.numPrims && numQueries >= kQueryOrderMin && (uint64_t)numQueries * k * 2 >= bvh.numPrims)
      CUBQL_CHECK(scr.alloc(&spts, bvh.numPrims));
  }
  if (k <= 4) return launch_knn<4>(bvh, d_points, d_queries, numQueries, k, sqrMaxSearchDist, d_outIDs, d_outSqrDist, d_numFound, ovf, ovfList, order, spts, s);
  if (k <= 8) return launch_knn<8>(bvh, d_points, d_queries, numQueries, k, sqrMaxSearchDist, d_outIDs, d_outSqrDist, d_numFound, ovf, ovfList, order, spts, s);
  if (k <= 16) return launch_knn<16>(bvh, d_points, d_queries, numQueries, k, sqrMaxSearchDist, d_outIDs, d_outSqrDist, d_numFound, ovf, ovfList, order, spts, s);
  if (k <= 32) return launch_knn<32>(bvh, d_points, d_queries, numQueries, k, sqrMaxSearchDist, d_outIDs, d_outSqrDist, d_numFound, ovf, ovfList, order, spts, s);
  return launch_knn<64>(bvh, d_points, d_queries, numQueries, k, sqrMaxSearchDist, d_outIDs, d_outSqrDist, d_numFound, ovf, ovfList, order, spts, s);
}

template <int W>
cubql_dev::WideBVH<W> wide_view(const cubql_wbvh3f &w) {
  return cubql_dev::WideBVH<W>{(const float4 *)w.nodes, w.numNodes, w.primIDs, w.numPrims};
}

}  // namespace pq
}  // namespace cubql

using namespace cubql;

extern "C" int cubql_fcp_points_f3(const cubql_bvh3f *bvh, const float *d_points,
                                   const float *d_queries, uint32_t numQueries,
                                   float sqrMaxSearchDist, int32_t *d_outIDs, float *d_outSqrDist,
                                   cubql_stream_t stream) {
  clear_error();
  if (!bvh || (!d_queries && numQueries) || (!d_outIDs && numQueries)) {
    set_error("cubql_fcp_points_f3: null argument");
    return CUBQL_ERR_INVALID_ARG;
  }
  if (numQueries == 0) return CUBQL_OK;
  return pq::fcp_run(*bvh, d_points, d_queries, numQueries, sqrMaxSearchDist, d_outIDs, d_outSqrDist,
                     (hipStream_t)stream, "fcp_points");
}

extern "C" int cubql_knn_points_f3(const cubql_bvh3f *bvh, const float *d_points,
                                   const float *d_queries, uint32_t numQueries, int k,
                                   float sqrMaxSearchDist, int32_t *d_outIDs, float *d_outSqrDist,
                                   int32_t *d_numFound, cubql_stream_t stream) {
  clear_error();
  if (!bvh || k < 1 || k > 64 || (numQueries && (!d_queries || !d_outIDs || !d_outSqrDist))) {
    set_error("cubql_knn_points_f3: invalid argument (k=%d)", k);
    return CUBQL_ERR_INVALID_ARG;
  }
  if (numQueries == 0) return CUBQL_OK;
  return pq::knn_run(*bvh, d_points, d_queries, numQueries, k, sqrMaxSearchDist, d_outIDs, d_outSqrDist,
                     d_numFound, (hipStream_t)stream);
}

extern "C" int cubql_fcp_points_wide_f3(const cubql_wbvh3f *bvh, int width, const float *d_points,
                                        const float *d_queries, uint32_t numQueries, float sqrMaxSearchDist,
                                        int32_t *d_outIDs, float *d_outSqrDist, cubql_stream_t stream) {
  clear_error();
  if (!bvh || (width != 4 && width != 8) || (!d_queries && numQueries) || (!d_outIDs && numQueries)) {
    set_error("cubql_fcp_points_wide_f3: invalid argument (width %d)", width);
    return CUBQL_ERR_INVALID_ARG;
  }
  if (numQueries == 0) return CUBQL_OK;
  hipStream_t s = (hipStream_t)stream;
  if (width == 4)
    return pq::fcp_run(pq::wide_view<4>(*bvh), d_points, d_queries, numQueries, sqrMaxSearchDist, d_outIDs,
                       d_outSqrDist, s, "fcp_points_wide4");
  return pq::fcp_run(pq::wide_view<8>(*bvh), d_points, d_queries, numQueries, sqrMaxSearchDist, d_outIDs,
                     d_outSqrDist, s, "fcp_points_wide8");
}

extern "C" int cubql_knn_points_wide_f3(const cubql_wbvh3f *bvh, int width, const float *d_points,
                                        const float *d_queries, uint32_t numQueries, int k, float sqrMaxSearchDist,
                                        int32_t *d_outIDs, float *d_outSqrDist, int32_t *d_numFound,
                                        cubql_stream_t stream) {
  clear_error();
  if (!bvh || (width != 4 && width != 8) || k < 1 || k > 64 ||
      (numQueries && (!d_queries || !d_outIDs || !d_outSqrDist))) {
    set_error("cubql_knn_points_wide_f3: invalid argument (width %d, k=%d)", width, k);
    return CUBQL_ERR_INVALID_ARG;
  }
  if (numQueries == 0) return CUBQL_OK;
  hipStream_t s = (hipStream_t)stream;
  if (width == 4)
    return pq::knn_run(pq::wide_view<4>(*bvh), d_points, d_queries, numQueries, k, sqrMaxSearchDist, d_outIDs,
                       d_outSqrDist, d_numFound, s);
  return pq::knn_run(pq::wide_view<8>(*bvh), d_points, d_queries, numQueries, k, sqrMaxSearchDist, d_outIDs,
                     d_outSqrDist, d_numFound, s);
}

// ------------------------------------------- closest points, no allocation
//
// cubql_fcp_points_noalloc_f3 runs cubql_fcp_points_f3's kernels on a
// workspace in a file-scope device array instead of allocated scratch, for
// callers that cannot allocate.  The array is shared by every call on a
// device and keeps the last call's contents; the overflow counters are
// cleared by a fill kernel at the start of each call.  Calls that overlap on
// one device (two streams, two threads) race on it.
namespace cubql {
namespace pq {
constexpr uint32_t kNoAllocMaxQueries = 1u << 20;
constexpr size_t kNoAllocScratchBytes = (size_t)24 << 20;
constexpr int kMaxDevices = 64;
}  // namespace pq
}  // namespace cubql

static __device__ __align__(256) char scratch[cubql::pq::kNoAllocScratchBytes];

namespace {
std::mutex g_scratchMu;
char *g_scratchPtr[cubql::pq::kMaxDevices];

// Device address of `scratch` on the current device, looked up on the
// first call there (the host-side name is not a device address).
int scratch_ptr(char **out) {
  using namespace cubql;
  int dev = -1;
  CUBQL_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= pq::kMaxDevices) {
    set_error("cubql_fcp_points_noalloc_f3: device %d out of range", dev);
    return CUBQL_ERR_INVALID_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratchMu);
  if (!g_scratchPtr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("cubql_fcp_points_noalloc_f3: hipGetSymbolAddress failed: %s", hipGetErrorString(e));
      return CUBQL_ERR_HIP;
    }
    g_scratchPtr[dev] = (char *)p;
  }
  *out = g_scratchPtr[dev];
  return CUBQL_OK;
}
}  // namespace

extern "C" int cubql_fcp_points_noalloc_f3(const cubql_bvh3f *bvh, const float *d_points, const float *d_queries,
                                           uint32_t numQueries, float sqrMaxSearchDist, int32_t *d_outIDs,
                                           float *d_outSqrDist, cubql_stream_t stream) {
  clear_error();
  if (!bvh || (!d_queries && numQueries) || (!d_outIDs && numQueries)) {
    set_error("cubql_fcp_points_noalloc_f3: null argument");
    return CUBQL_ERR_INVALID_ARG;
  }
  if (numQueries > pq::kNoAllocMaxQueries) {
    set_error("cubql_fcp_points_noalloc_f3: %u queries exceed the limit of %u per call", numQueries,
              pq::kNoAllocMaxQueries);
    return CUBQL_ERR_INVALID_ARG;
  }
  if (numQueries == 0) return CUBQL_OK;
  if (pq::fcp_ws_bytes(numQueries) > pq::kNoAllocScratchBytes) {
    set_error("cubql_fcp_points_noalloc_f3: workspace of %zu bytes exceeds the static %zu", pq::fcp_ws_bytes(numQueries),
              pq::kNoAllocScratchBytes);
    return CUBQL_ERR_INTERNAL;
  }
  char *ws = nullptr;
  CUBQL_CHECK(scratch_ptr(&ws));
  return pq::fcp_run_ws(*bvh, d_points, d_queries, numQueries, sqrMaxSearchDist, d_outIDs, d_outSqrDist,
                        (hipStream_t)stream, ws, "fcp_points_noalloc");
}

// ------------------------------------------------- traversal statistics
namespace cubql {
namespace pq {

__device__ __forceinline__ void stats_flush(const cubql_dev::TraversalCounts &st, uint32_t nq,
                                            unsigned long long *out) {
  unsigned long long nb = st.nodeBytes, np = st.prims;
  for (int o = 32; o > 0; o >>= 1) {
    nb += __shfl_xor(nb, o, 64);
    np += __shfl_xor(np, o, 64);
  }
  if (threadIdx.x == 0) {
    atomicAdd(&out[0], nb);
    atomicAdd(&out[1], np);
    atomicAdd(&out[2], (unsigned long long)min(64u, nq - blockIdx.x * 64));
  }
}

// Instrumented closest-point (k == 0) or kNN queries with the deep stack:
// per query, bytes of node data loaded and prims tested.
template <int K>
__global__ __launch_bounds__(64) void k_points_stats(cubql_bvh3f bvh, const float *points, const float *queries,
                                                     uint32_t nq, int k, float sqrMax, unsigned long long *out) {
  const uint32_t qi = blockIdx.x * 64 + threadIdx.x;
  cubql_dev::TraversalCounts st;
  if (qi < nq) {
    const float3 q = load_pt(queries, qi);
    if (k == 0) {
      int32_t best;
      float bestD;
      fcp_one<kDeepStack>(bvh, points, q, sqrMax, best, bestD, st);
    } else {
      int32_t ids[K];
      float ds[K];
      int found;
      // the traversal k_knn runs, on the deep stack
      if constexpr (K <= kKnnSeededMax) {
        cubql_dev::PrivateStack<kDeepStack> stack;
        knn_seeded_on<K>(stack, bvh, points, nullptr, q, k, sqrMax, ids, ds, found, st);
      } else {
        knn_one<K, kDeepStack>(bvh, points, nullptr, q, k, sqrMax, ids, ds, found, st);
      }
    }
  }
  stats_flush(st, nq, out);
}

}  // namespace pq
}  // namespace cubql

extern "C" int cubql_query_stats_points_f3(const cubql_bvh3f *bvh, const float *d_points, const float *d_queries,
                                           uint32_t numQueries, int k, float sqrMaxSearchDist, uint64_t *d_out,
                                           cubql_stream_t stream) {
  using namespace cubql;
  clear_error();
  if (!bvh || !d_out || k < 0 || k > 64 || (numQueries && (!d_queries || !d_points))) {
    set_error("cubql_query_stats_points_f3: invalid argument (k=%d)", k);
    return CUBQL_ERR_INVALID_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  CUBQL_CHECK(fill_u32((uint32_t *)d_out, 0u, 6, s));
  if (numQueries == 0 || bvh->numNodes == 0) return CUBQL_OK;
  const uint32_t g = div_up(numQueries, 64);
  unsigned long long *o = (unsigned long long *)d_out;
  if (k <= 4)
    pq::k_points_stats<4><<<g, 64, 0, s>>>(*bvh, d_points, d_queries, numQueries, k, sqrMaxSearchDist, o);
  else if (k <= 8)
    pq::k_points_stats<8><<<g, 64, 0, s>>>(*bvh, d_points, d_queries, numQueries, k, sqrMaxSearchDist, o);
  else if (k <= 16)
    pq::k_points_stats<16><<<g, 64, 0, s>>>(*bvh, d_points, d_queries, numQueries, k, sqrMaxSearchDist, o);
  else if (k <= 32)
    pq::k_points_stats<32><<<g, 64, 0, s>>>(*bvh, d_points, d_queries, numQueries, k, sqrMaxSearchDist, o);
  else
    pq::k_points_stats<64><<<g, 64, 0, s>>>(*bvh, d_points, d_queries, numQueries, k, sqrMaxSearchDist, o);
  CUBQL_LAUNCH_CHECK();
  return CUBQL_OK;
}
