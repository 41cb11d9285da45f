// triangles.hip -- closest point on a triangle mesh, and distance adjoints.
//
// Forward: replaces cuBQL::triangles::findClosest / closestPoint
// (queries/triangles/findClosest.h:56-152) as a batched kernel, one query
// per lane, over an indexed mesh (tris[3i..3i+2] index verts) or, when tris
// is NULL, the reference's non-indexed Triangle array (verts[9i..9i+8]).
// The per-triangle test is the reference's: N = (b-a)x(c-a); outside if
// N == 0 or q lies beyond any edge plane, then the closest of the three
// clamped edge projections; else the plane distance along
// normalize(N) = N*(1/sqrt(N.N)).  Float32 with contraction off; among equal
// distances the lowest triangle id wins.
//
// Backward (no reference counterpart): recomputes the closest feature of
// the saved triangle id -- no traversal state is stored -- and applies the
// envelope theorem: for d = |q - P|, P = sum_j w_j v_j,
//   dd/dq = (q-P)/d,  dd/dv_j = -w_j (q-P)/d,   (0 at d == 0)
// with w the barycentric weights of P (face: projection barycentrics,
// edge: (1-t, t), vertex: one-hot).  Vertex gradients are scatter-added.
#include "common.h"
#include "cubql_amd/traversal.hip.h"
#include "cubql_amd/triangle.hip.h"
#include "profile.h"
#include "query_order.h"
#include "radix_sort.h"

namespace cubql {
namespace tri {

constexpr int kBlock = 256;
constexpr int kFastStack = 64;
constexpr int kDeepStack = 1024;
// id of a query whose traversal needed more than kDeepStack entries
constexpr int32_t kIncomplete = -2;

using namespace cubql_dev;

struct Overflow { uint32_t count, deepFailed; };

template <int STACK, typename Stats = NoStats>
__device__ __forceinline__ bool fcp_tri_one(const cubql_bvh3f &bvh, const int32_t *tris, const float *verts,
                                            F3 q, float sqrMax, int32_t &best, float &bestD, F3 &bestP,
                                            Stats &&st = Stats()) {
  best = -1;
  bestD = sqrMax;
  bestP = {0.f, 0.f, 0.f};
  auto prim = [&](uint32_t t) -> float {
    st.prim();
    F3 a, b, c, P;
    tri_verts(tris, verts, t, a, b, c);
    float w[3];
    const float d = tri_closest<false>(a, b, c, q, P, w);
    if (d < bestD || (d == bestD && best >= 0 && (int32_t)t < best)) {
      bestD = d;
      best = (int32_t)t;
      bestP = P;
    }
    return bestD;
  };
  // seed the radius with the nearest-child leaf, which the traversal then
  // leaves out
  const uint32_t seed = cubql_dev::visit_nearest_leaf(bvh, q.x, q.y, q.z, prim, 0u, st);
  return cubql_dev::for_each_prim_except<STACK>(bvh, q.x, q.y, q.z, bestD, prim, seed, 0u, st);
}

// The same on a caller-provided traversal stack.
template <typename StackT>
__device__ __forceinline__ bool fcp_tri_one_on(StackT &stack, const cubql_bvh3f &bvh, const int32_t *tris,
                                               const float *verts, F3 q, float sqrMax, int32_t &best, float &bestD,
                                               F3 &bestP) {
  best = -1;
  bestD = sqrMax;
  bestP = {0.f, 0.f, 0.f};
  auto prim = [&](uint32_t t) -> float {
    F3 a, b, c, P;
    tri_verts(tris, verts, t, a, b, c);
    float w[3];
    const float d = tri_closest<false>(a, b, c, q, P, w);
    if (d < bestD || (d == bestD && best >= 0 && (int32_t)t < best)) {
      bestD = d;
      best = (int32_t)t;
      bestP = P;
    }
    return bestD;
  };
  NoStats st;
  const uint32_t seed = cubql_dev::visit_nearest_leaf(bvh, q.x, q.y, q.z, prim, 0u, st);
  return cubql_dev::for_each_prim_except_on(stack, bvh, q.x, q.y, q.z, bestD, prim, seed, 0u, st);
}


// The same with the leaf triangles read from a copy of the mesh in
// primID-slot order (three float4 per slot: a and the triangle id's bits, b,
// c): a leaf's triangles are one contiguous run instead of a primID, index
// and vertex gather chain each.  Tests the same triangles in the same order
// with the same vertices, so the result is the same.
template <typename StackT>
__device__ __forceinline__ bool fcp_tri_sorted_on(StackT &stack, const cubql_bvh3f &bvh, const float4 *stris, F3 q,
                                                  float sqrMax, int32_t &best, float &bestD, F3 &bestP) {
  best = -1;
  bestD = sqrMax;
  bestP = {0.f, 0.f, 0.f};
  auto test = [&](uint32_t slot) {
    const float4 va = stris[3 * (size_t)slot], vb = stris[3 * (size_t)slot + 1], vc = stris[3 * (size_t)slot + 2];
    const F3 a = {va.x, va.y, va.z}, b = {vb.x, vb.y, vb.z}, c = {vc.x, vc.y, vc.z};
    const int32_t t = (int32_t)__float_as_uint(va.w);
    F3 P;
    float w[3];
    const float d = tri_closest<false>(a, b, c, q, P, w);
    if (d < bestD || (d == bestD && best >= 0 && t < best)) {
      bestD = d;
      best = t;
      bestP = P;
    }
  };
  NoStats st;
  const uint64_t seedLeaf = nearest_leaf(bvh, q.x, q.y, q.z, 0u, st);
  const uint32_t seed = oc_off(seedLeaf);
  for (uint32_t j = 0; j < oc_cnt(seedLeaf); j++) test(seed + j);
  auto leaf = [&](const uint32_t *ids, uint32_t cnt) -> float {
    const uint32_t off = (uint32_t)(ids - bvh.primIDs);
    if (off == seed) return INFINITY;
    for (uint32_t j = 0; j < cnt; j++) test(off + j);
    return bestD;
  };
  return for_each_leaf_on(stack, bvh, q.x, q.y, q.z, bestD, leaf, 0u, st);
}

// the mesh in primID-slot order for fcp_tri_sorted_on
__global__ __launch_bounds__(256) void k_slot_tris(const uint32_t *primIDs, uint32_t numPrims, const int32_t *tris,
                                                   const float *verts, float4 *stris) {
  for (uint32_t s = blockIdx.x * 256 + threadIdx.x; s < numPrims; s += gridDim.x * 256) {
    const uint32_t t = primIDs[s];
    F3 a, b, c;
    tri_verts(tris, verts, t, a, b, c);
    stris[3 * (size_t)s] = make_float4(a.x, a.y, a.z, __uint_as_float(t));
    stris[3 * (size_t)s + 1] = make_float4(b.x, b.y, b.z, 0.f);
    stris[3 * (size_t)s + 2] = make_float4(c.x, c.y, c.z, 0.f);
  }
}

__device__ __forceinline__ void write_result(uint32_t qi, int32_t best, float bestD, F3 P, int32_t *ids,
                                             float *d2, float *cp) {
  ids[qi] = best;
  if (d2) d2[qi] = best >= 0 ? bestD : INFINITY;
  if (cp) {
    cp[3 * (size_t)qi] = P.x;
    cp[3 * (size_t)qi + 1] = P.y;
    cp[3 * (size_t)qi + 2] = P.z;
  }
}

// kSorted: the leaves read stris, the mesh in primID-slot order
template <bool kSorted>
__global__ __launch_bounds__(kBlock) void k_fcp_tris(cubql_bvh3f bvh, const int32_t *tris, const float *verts,
                                                     const float4 *stris, const float *queries, uint32_t nq,
                                                     float sqrMax,
                                                     int32_t *ids, float *d2, float *cp, Overflow *ovf,
                                                     uint32_t *ovfList, const uint32_t *order) {
  // Morton-ordered queries: one contiguous run of them per XCD
  const uint32_t i = (order ? xcd_chunk(blockIdx.x, gridDim.x) : blockIdx.x) * kBlock + threadIdx.x;
  if (i >= nq) return;
  const uint32_t qi = order ? order[i] : i;
  const F3 q = ld3(queries, qi);
  int32_t best;
  float bestD;
  F3 P;
  // private stack: triangle queries on an SAH tree stack many far
  // children, and a short LDS stack sent too many of them to the deep rerun
  PrivateStack<kFastStack> stack;
  bool done;
  if constexpr (kSorted) done = fcp_tri_sorted_on(stack, bvh, stris, q, sqrMax, best, bestD, P);
  else done = fcp_tri_one_on(stack, bvh, tris, verts, q, sqrMax, best, bestD, P);
  if (!done) {
    ovfList[atomicAdd(&ovf->count, 1u)] = qi;
    return;
  }
  write_result(qi, best, bestD, P, ids, d2, cp);
}

__global__ __launch_bounds__(64) void k_fcp_tris_deep(cubql_bvh3f bvh, const int32_t *tris, const float *verts,
                                                      const float *queries, float sqrMax, int32_t *ids,
                                                      float *d2, float *cp, Overflow *ovf,
                                                      const uint32_t *ovfList) {
  const uint32_t n = ovf->count;
  for (uint32_t i = blockIdx.x * 64 + threadIdx.x; i < n; i += gridDim.x * 64) {
    const uint32_t qi = ovfList[i];
    const F3 q = ld3(queries, qi);
    int32_t best;
    float bestD;
    F3 P;
    if (!fcp_tri_one<kDeepStack>(bvh, tris, verts, q, sqrMax, best, bestD, P)) {
      // deeper than the deep stack: marked incomplete, not "none found"
      atomicAdd(&ovf->deepFailed, 1u);
      best = kIncomplete;
    }
    write_result(qi, best, bestD, P, ids, d2, cp);
  }
}

__global__ void k_tri_boxes(const int32_t *tris, const float *verts, uint32_t n, cubql_box3f *boxes) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    F3 a, b, c;
    tri_verts(tris, verts, i, a, b, c);
    cubql_box3f bx;  // Triangle::bounds() (Triangle.h:35-36)
    bx.lower[0] = fminf(fminf(a.x, b.x), c.x);
    bx.lower[1] = fminf(fminf(a.y, b.y), c.y);
    bx.lower[2] = fminf(fminf(a.z, b.z), c.z);
    bx.upper[0] = fmaxf(fmaxf(a.x, b.x), c.x);
    bx.upper[1] = fmaxf(fmaxf(a.y, b.y), c.y);
    bx.upper[2] = fmaxf(fmaxf(a.z, b.z), c.z);
    boxes[i] = bx;
  }
}

__global__ __launch_bounds__(kBlock) void k_tri_backward(const int32_t *tris, const float *verts,
                                                         const float *queries, uint32_t nq,
                                                         const int32_t *ids, const float *gradDist,
                                                         float *gradQ, float *gradV) {
  const uint32_t qi = blockIdx.x * kBlock + threadIdx.x;
  if (qi >= nq) return;
  const int32_t t = ids[qi];
  const F3 q = ld3(queries, qi);
  F3 gq = {0.f, 0.f, 0.f};
  if (t >= 0) {
    F3 a, b, c, P;
    tri_verts(tris, verts, (uint32_t)t, a, b, c);
    float w[3];
    const float s = tri_closest<true>(a, b, c, q, P, w);
    const float d = sqrtf(s);
    if (d > 0.f) {
      const float g = gradDist[qi] / d;
      const F3 dir = sub(q, P);
      gq = mul(g, dir);
      if (gradV) {
        for (int j = 0; j < 3; j++) {
          const size_t v = tris ? (size_t)tris[3 * (size_t)t + j] : 3 * (size_t)t + j;
          atomicAdd(&gradV[3 * v + 0], -w[j] * gq.x);
          atomicAdd(&gradV[3 * v + 1], -w[j] * gq.y);
          atomicAdd(&gradV[3 * v + 2], -w[j] * gq.z);
        }
      }
    }
  }
  if (gradQ) {
    gradQ[3 * (size_t)qi] = gq.x;
    gradQ[3 * (size_t)qi + 1] = gq.y;
    gradQ[3 * (size_t)qi + 2] = gq.z;
  }
}

// points: d = |q-p| for each (query, neighbour) entry e = qi*k + j.  The
// query gradient is summed per lane.  The point gradient is not scattered
// here: when gv is given, each entry's point term -g (0 for skipped
// entries) goes to gv[e] and its sort key to key[e] (the point id, or
// numPoints for entries that contribute nothing), val[e] = e.
__global__ __launch_bounds__(kBlock) void k_point_backward(const float *points, const float *queries,
                                                           uint32_t nq, int k, const int32_t *ids,
                                                           const float *gradDist, float *gradQ, float *gv,
                                                           uint32_t *key, uint32_t *val, uint32_t numPoints) {
  const uint32_t qi = blockIdx.x * kBlock + threadIdx.x;
  if (qi >= nq) return;
  const F3 q = ld3(queries, qi);
  F3 gq = {0.f, 0.f, 0.f};
  for (int j = 0; j < k; j++) {
    const size_t e = (size_t)qi * k + j;
    const int32_t p = ids[e];
    F3 g = {0.f, 0.f, 0.f};
    bool live = false;
    if (p >= 0) {
      const F3 diff = sub(q, ld3(points, (size_t)p));
      const float d = sqrtf(dot(diff, diff));
      if (d > 0.f) {
        g = mul(gradDist[e] / d, diff);
        gq = add(gq, g);
        live = true;
      }
    }
    if (gv) {
      gv[3 * e + 0] = -g.x;
      gv[3 * e + 1] = -g.y;
      gv[3 * e + 2] = -g.z;
      key[e] = live && (uint32_t)p < numPoints ? (uint32_t)p : numPoints;
      val[e] = (uint32_t)e;
    }
  }
  if (gradQ) {
    gradQ[3 * (size_t)qi] = gq.x;
    gradQ[3 * (size_t)qi + 1] = gq.y;
    gradQ[3 * (size_t)qi + 2] = gq.z;
  }
}

// The point gradient from the entries sorted by point id: each wave sums
// its runs of equal ids (contiguous after the sort) with a segmented scan;
// a run inside the wave holds all of its point's entries, so it is added
// with a plain read-modify-write, and only runs that touch the wave's first
// or last lane (which may continue in the next or previous wave) use
// atomics.  No float atomic per entry, and the sum order inside a run is
// fixed.
__global__ __launch_bounds__(kBlock) void k_point_scatter(const uint32_t *key, const uint32_t *val,
                                                          const float *gv, uint32_t n, uint32_t numPoints,
                                                          float *gradP) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t lane = lane_id();
  const bool act = i < n;
  const uint32_t p = act ? key[i] : numPoints;
  float sx = 0.f, sy = 0.f, sz = 0.f;
  if (act && p < numPoints) {
    const uint32_t e = val[i];
    sx = gv[3 * (size_t)e + 0];
    sy = gv[3 * (size_t)e + 1];
    sz = gv[3 * (size_t)e + 2];
  }
  const uint32_t pPrev = (uint32_t)__shfl_up((int)p, 1, kWave);
  const bool head = lane == 0 || pPrev != p;
  bool seen = head;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const float ux = __shfl_up(sx, o, kWave), uy = __shfl_up(sy, o, kWave), uz = __shfl_up(sz, o, kWave);
    const bool us = __shfl_up((int)seen, o, kWave) != 0;
    if (lane >= (uint32_t)o && !seen) {
      sx += ux; sy += uy; sz += uz;
      seen = us;
    }
  }
  const uint32_t pNext = (uint32_t)__shfl_down((int)p, 1, kWave);
  const bool runEnd = lane == kWave - 1 || pNext != p;
  // lane 0 of a run: the lane where the run starts (a ballot over heads)
  const uint64_t heads = __ballot(head);
  const uint64_t below = heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1));
  const uint32_t runStart = 63u - (uint32_t)__builtin_clzll(below);
  if (runEnd && p < numPoints) {
    float *o = gradP + 3 * (size_t)p;
    if (runStart == 0 || lane == kWave - 1) {
      atomicAdd(&o[0], sx);
      atomicAdd(&o[1], sy);
      atomicAdd(&o[2], sz);
    } else {
      o[0] += sx;
      o[1] += sy;
      o[2] += sz;
    }
  }
}

__global__ void k_fill_nf(uint32_t nq, int32_t *ids, float *d2) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x) {
    ids[i] = -1;
    if (d2) d2[i] = INFINITY;
  }
}

}  // namespace tri
}  // namespace cubql

using namespace cubql;

extern "C" int cubql_triangle_boxes_f3(const int32_t *d_tris, const float *d_verts, uint32_t numTris,
                                       cubql_box3f *d_boxes, cubql_stream_t s) {
  clear_error();
  if (numTris == 0) return CUBQL_OK;
  if (!d_verts || !d_boxes) {
    set_error("cubql_triangle_boxes_f3: null argument");
    return CUBQL_ERR_INVALID_ARG;
  }
  tri::k_tri_boxes<<<std::min<uint32_t>(div_up(numTris, 256), 8192), 256, 0, (hipStream_t)s>>>(
      d_tris, d_verts, numTris, d_boxes);
  CUBQL_LAUNCH_CHECK();
  return CUBQL_OK;
}

extern "C" int cubql_fcp_triangles_f3(const cubql_bvh3f *bvh, const int32_t *d_tris, const float *d_verts,
                                      const float *d_queries, uint32_t numQueries, float sqrMax,
                                      int32_t *d_outTriIDs, float *d_outSqrDist, float *d_outClosest,
                                      cubql_stream_t stream) {
  clear_error();
  if (!bvh || (numQueries && (!d_verts || !d_queries || !d_outTriIDs))) {
    set_error("cubql_fcp_triangles_f3: null argument");
    return CUBQL_ERR_INVALID_ARG;
  }
  if (numQueries == 0) return CUBQL_OK;
  hipStream_t s = (hipStream_t)stream;
  if (bvh->numNodes == 0) {
    tri::k_fill_nf<<<256, 256, 0, s>>>(numQueries, d_outTriIDs, d_outSqrDist);
    CUBQL_LAUNCH_CHECK();
    return CUBQL_OK;
  }
  Scratch scr(nullptr, s);
  tri::Overflow *ovf = nullptr;
  uint32_t *ovfList = nullptr;
  CUBQL_CHECK(scr.alloc(&ovf, 1));
  CUBQL_CHECK(scr.alloc(&ovfList, numQueries));
  CUBQL_CHECK(fill_u32((uint32_t *)ovf, 0u, 2, s));
  const uint32_t *order = nullptr;
  CUBQL_CHECK(query_order(*bvh, d_queries, numQueries, s, scr, &order));
  // the slot-ordered copy of the mesh pays when the queries test at least
  // about as many triangles as the copy moves
  float4 *stris = nullptr;
  if (bvh->numPrims && numQueries >= kQueryOrderMin && (uint64_t)numQueries * 8 >= bvh->numPrims)
    CUBQL_CHECK(scr.alloc(&stris, 3 * (size_t)bvh->numPrims));
  int ps = prof_begin(s);
  if (stris) {
    tri::k_slot_tris<<<std::max(1u, std::min(div_up(bvh->numPrims, 256u), 2048u)), 256, 0, s>>>(
        bvh->primIDs, bvh->numPrims, d_tris, d_verts, stris);
    tri::k_fcp_tris<true><<<div_up(numQueries, tri::kBlock), tri::kBlock, 0, s>>>(
        *bvh, d_tris, d_verts, stris, d_queries, numQueries, sqrMax, d_outTriIDs, d_outSqrDist, d_outClosest, ovf,
        ovfList, order);
  } else {
    tri::k_fcp_tris<false><<<div_up(numQueries, tri::kBlock), tri::kBlock, 0, s>>>(
        *bvh, d_tris, d_verts, nullptr, d_queries, numQueries, sqrMax, d_outTriIDs, d_outSqrDist, d_outClosest,
        ovf, ovfList, order);
  }
  prof_end(ps, s, "fcp_triangles");
  prof_set(ps, 20.0 * numQueries, numQueries);
  tri::k_fcp_tris_deep<<<64, 64, 0, s>>>(*bvh, d_tris, d_verts, d_queries, sqrMax, d_outTriIDs,
                                         d_outSqrDist, d_outClosest, ovf, ovfList);
  CUBQL_LAUNCH_CHECK();
  return CUBQL_OK;
}

extern "C" int cubql_fcp_triangles_backward_f3(const int32_t *d_tris, const float *d_verts,
                                               const float *d_queries, uint32_t numQueries,
                                               const int32_t *d_triIDs, const float *d_gradDist,
                                               float *d_gradQueries, float *d_gradVerts,
                                               cubql_stream_t s) {
  clear_error();
  if (numQueries == 0) return CUBQL_OK;
  if (!d_verts || !d_queries || !d_triIDs || !d_gradDist) {
    set_error("cubql_fcp_triangles_backward_f3: null argument");
    return CUBQL_ERR_INVALID_ARG;
  }
  int ps = prof_begin((hipStream_t)s);
  tri::k_tri_backward<<<div_up(numQueries, tri::kBlock), tri::kBlock, 0, (hipStream_t)s>>>(
      d_tris, d_verts, d_queries, numQueries, d_triIDs, d_gradDist, d_gradQueries, d_gradVerts);
  prof_end(ps, (hipStream_t)s, "triangles_backward");
  prof_set(ps, (12.0 + 4.0 + 4.0 + 12.0 + 3 * 12.0 + 3 * 24.0) * numQueries, numQueries);
  CUBQL_LAUNCH_CHECK();
  return CUBQL_OK;
}

extern "C" int cubql_points_distance_backward_f3(const float *d_points, uint32_t numPoints, const float *d_queries,
                                                 uint32_t numQueries, int k, const int32_t *d_ids,
                                                 const float *d_gradDist, float *d_gradQueries,
                                                 float *d_gradPoints, cubql_stream_t stream,
                                                 const cubql_mem_resource *mem) {
  clear_error();
  if (numQueries == 0) return CUBQL_OK;
  if (!d_points || !d_queries || !d_ids || !d_gradDist || k < 1) {
    set_error("cubql_points_distance_backward_f3: invalid argument");
    return CUBQL_ERR_INVALID_ARG;
  }
  const size_t ne = (size_t)numQueries * (size_t)k;
  if (ne > 0xffffffffull || numPoints > 0x7fffffffu) {
    set_error("cubql_points_distance_backward_f3: %zu entries / %u points exceed the 32-bit entry index", ne,
              numPoints);
    return CUBQL_ERR_INVALID_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const bool scatter = d_gradPoints && numPoints;
  ScratchArena scr(mem, s);
  float *gv = nullptr;
  uint32_t *kA = nullptr, *vA = nullptr, *kB = nullptr, *vB = nullptr, *cnt = nullptr;
  if (scatter) {
    CUBQL_CHECK(scr.reserve(&gv, 3 * ne));
    CUBQL_CHECK(scr.reserve(&kA, ne));
    CUBQL_CHECK(scr.reserve(&vA, ne));
    CUBQL_CHECK(scr.reserve(&kB, ne));
    CUBQL_CHECK(scr.reserve(&vB, ne));
    CUBQL_CHECK(scr.reserve(&cnt, 1));
    CUBQL_CHECK(scr.commit());
    CUBQL_CHECK(fill_u32(cnt, (uint32_t)ne, 1, s));
  }
  int ps = prof_begin(s);
  tri::k_point_backward<<<div_up(numQueries, tri::kBlock), tri::kBlock, 0, s>>>(
      d_points, d_queries, numQueries, k, d_ids, d_gradDist, d_gradQueries, gv, kA, vA, numPoints);
  CUBQL_LAUNCH_CHECK();
  if (scatter) {
    // key bits: enough for the sentinel numPoints
    int bits = 1;
    while (bits < 32 && (numPoints >> bits) != 0) bits++;
    uint32_t *kRes = nullptr, *vRes = nullptr;
    CUBQL_CHECK(radix_sort_pairs_dev_u32(kA, vA, kB, vB, cnt, (uint32_t)ne, bits, s, mem, &kRes, &vRes));
    tri::k_point_scatter<<<div_up((uint32_t)ne, tri::kBlock), tri::kBlock, 0, s>>>(kRes, vRes, gv, (uint32_t)ne,
                                                                                  numPoints, d_gradPoints);
    CUBQL_LAUNCH_CHECK();
  }
  prof_end(ps, s, "points_backward");
  prof_set(ps, (12.0 + 12.0 + k * (4.0 + 4.0 + 12.0 + 24.0)) * numQueries, numQueries);
  return CUBQL_OK;
}

// ------------------------------------------------- traversal statistics
namespace cubql {
namespace tri {

// Instrumented closest-triangle queries (deep stack): per query, bytes of
// node data loaded and triangles tested, summed into out[0], out[1];
// out[2] counts the queries.
__global__ __launch_bounds__(64) void k_fcp_tris_stats(cubql_bvh3f bvh, const int32_t *tris, const float *verts,
                                                       const float *queries, uint32_t nq, float sqrMax,
                                                       unsigned long long *out) {
  const uint32_t qi = blockIdx.x * 64 + threadIdx.x;
  TraversalCounts st;
  if (qi < nq) {
    int32_t best;
    float bestD;
    F3 P;
    fcp_tri_one<kDeepStack>(bvh, tris, verts, ld3(queries, qi), sqrMax, best, bestD, P, st);
  }
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

}  // namespace tri
}  // namespace cubql

extern "C" int cubql_query_stats_triangles_f3(const cubql_bvh3f *bvh, const int32_t *d_tris, const float *d_verts,
                                              const float *d_queries, uint32_t numQueries, float sqrMaxSearchDist,
                                              uint64_t *d_out, cubql_stream_t stream) {
  using namespace cubql;
  clear_error();
  if (!bvh || !d_verts || !d_out || (numQueries && !d_queries)) {
    set_error("cubql_query_stats_triangles_f3: null argument");
    return CUBQL_ERR_INVALID_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  CUBQL_CHECK(fill_u32((uint32_t *)d_out, 0u, 6, s));
  if (numQueries == 0 || bvh->numNodes == 0) return CUBQL_OK;
  tri::k_fcp_tris_stats<<<div_up(numQueries, 64), 64, 0, s>>>(*bvh, d_tris, d_verts, d_queries, numQueries,
                                                              sqrMaxSearchDist, (unsigned long long *)d_out);
  CUBQL_LAUNCH_CHECK();
  return CUBQL_OK;
}
