// batch.hip -- batched build + bidirectional closest point + Chamfer adjoint
// over thousands of independent point-cloud pairs (no reference counterpart:
// the reference builds one BVH per call and syncs with the host every level).
//
// Build, for each side (source / target) at once:
//   * one workgroup per cloud computes the cloud's centroid bounds and its
//     30-bit Morton keys with that cloud's own quantizer (radix.h:41-70); the
//     sort key is (cloud << 30) | morton, the value the global point index;
//   * one stable onesweep radix sort over all keys (segmented Morton sort);
//   * one level-synchronous top-down hierarchy over the sorted keys where a
//     range that spans several clouds is always split (its highest differing
//     key bit is a cloud bit, so the split lands on a cloud boundary) and a
//     single-cloud range follows radix.h:446-570 (leaf if <= T, equal keys
//     split in the middle).  The node of each single-cloud range is that
//     cloud's root, and its subtree is exactly the tree radix30Builder builds
//     for the cloud alone;
//   * exact bottom-up refit.
// Queries run one per lane over all clouds, in Morton order of the query
// side, each from its cloud's root in the other side's forest; the adjoint of
// loss[c] = mean_i d(s_i, T_c) + mean_j d(t_j, S_c) is scatter-added.  The
// whole sequence is a fixed number of launches on one stream with no host
// synchronisation and no allocation (caller workspace), so it can be
// captured in a HIP graph and replayed.
#include "block_ops.h"
#include "build_common.h"
#include "common.h"
#include "cubql_amd/traversal.hip.h"
#include "profile.h"
#include "radix_sort.h"

#include <algorithm>
#include <vector>

namespace cubql {
namespace bt {

constexpr int kCloudShift = 30;
constexpr uint32_t kNoRoot = 0xffffffffu;
constexpr uint32_t kNoSlot = 0xffffffffu;

struct Side {
  const float *pts;
  const uint32_t *off;  // numClouds + 1 offsets (device)
  uint32_t n;
  // persistent (kept for the queries)
  uint64_t *keys;       // sorted keys
  uint32_t *ids;        // sorted global point indices = primIDs of the forest
  float4 *spts;         // points in sorted order: x, y, z, global index (bits)
  uint32_t *nnS;        // per sorted query: sorted slot of its closest point
  float *d2S;           //   on the other side (kNoSlot if none) and its distance
  Node *nodes;          // capacity 2n + 2
  uint32_t *rootOf;     // numClouds
  uint32_t *ctr;        // counters, see C_* (C_WORDS words)
};

struct Temp {
  uint64_t *keysB;
  uint32_t *valsB;
  uint2 *ranges;
  uint32_t *levels;     // node ids level by level (capacity 2n+2)
  void *sortWs;
  uint32_t *ovf;        // [0] count, [1] deep failures, then the list
};

// [C_LEVEL0 + l] node count of level l; [C_LSTART0 + l] its start in the
// all-levels node list
enum { C_NODES = 0, C_N = 1, C_LEVEL0 = 8, C_LSTART0 = 256, C_WORDS = 512 };
constexpr int kMaxLevels = C_LSTART0 - C_LEVEL0 - 4;

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  uint32_t x = v & 0x3ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

// one workgroup per cloud: centroid bounds, quantizer, keys
__global__ __launch_bounds__(256) void k_bkeys(Side sd, uint32_t numClouds) {
  __shared__ float sb[12 * 4];
  const uint32_t c = blockIdx.x;
  const uint32_t b = sd.off[c], e = sd.off[c + 1];
  float v[12];
  for (int j = 0; j < 3; j++) { v[j] = INFINITY; v[j + 3] = -INFINITY; v[j + 6] = INFINITY; v[j + 9] = -INFINITY; }
  for (uint32_t i = b + threadIdx.x; i < e; i += 256) {
    for (int d = 0; d < 3; d++) {
      const float p = sd.pts[3 * (size_t)i + d];
      const float ce = 0.5f * (p + p);  // centre of the degenerate box (p, p)
      v[d] = fminf(v[d], ce);
      v[d + 3] = fmaxf(v[d + 3], ce);
    }
  }
  block_reduce_bounds<256>(v, sb);
  if (b == e) return;
  // Quantizer::init (radix.h:46-53), 10 bits per axis
  const float ext = fmaxf(fmaxf(v[3] - v[0], v[4] - v[1]), v[5] - v[2]);
  const float scale = 1024.f * (1.f / fmaxf(ext, 1e-20f));
  for (uint32_t i = b + threadIdx.x; i < e; i += 256) {
    uint32_t cell[3];
    for (int d = 0; d < 3; d++) {
      const float p = sd.pts[3 * (size_t)i + d];
      const float f = (0.5f * (p + p) - v[d]) * scale;
      cell[d] = min((uint32_t)f, 1023u);
    }
    const uint32_t m = spread10(cell[0]) | (spread10(cell[1]) << 1) | (spread10(cell[2]) << 2);
    sd.keys[i] = ((uint64_t)c << kCloudShift) | m;
    sd.ids[i] = i;
  }
}

// points in sorted order, with their global index in w: a leaf's prims are
// then one contiguous run for the refit and the queries
__global__ __launch_bounds__(256) void k_bgather(Side sd) {
  for (uint32_t t = blockIdx.x * 256 + threadIdx.x; t < sd.n; t += gridDim.x * 256) {
    const uint32_t i = sd.ids[t];
    sd.spts[t] = make_float4(sd.pts[3 * (size_t)i], sd.pts[3 * (size_t)i + 1], sd.pts[3 * (size_t)i + 2],
                             __uint_as_float(i));
  }
}

__global__ void k_binit(Side sd, Temp tp, uint32_t numClouds) {
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < numClouds; c += gridDim.x * blockDim.x)
    sd.rootOf[c] = kNoRoot;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sd.ctr[C_NODES] = 2;
    Node z;
    for (int d = 0; d < 3; d++) { z.lo[d] = 0.f; z.hi[d] = 0.f; }
    z.oc = 0;
    sd.nodes[1] = z;
    sd.ctr[C_LSTART0] = 0;
    if (sd.n > 0) {
      tp.ranges[0] = make_uint2(0u, sd.n);
      tp.levels[0] = 0;
      sd.ctr[C_LEVEL0] = 1;
    }
  }
}

// the root of a forest that holds a single cloud is node 0
__global__ void k_bsingle_root(Side sd) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && sd.n > 0) {
    const uint64_t c0 = sd.keys[0] >> kCloudShift, c1 = sd.keys[sd.n - 1] >> kCloudShift;
    if (c0 == c1) sd.rootOf[c0] = 0;
  }
}

__device__ __forceinline__ uint32_t find_split64(const uint64_t *keys, uint32_t b, uint32_t e) {
  const uint64_t fk = keys[b], lk = keys[e - 1];
  if (fk == lk) return (uint32_t)(((uint64_t)b + e) / 2);
  const int nm = __clzll((unsigned long long)(fk ^ lk));
  const uint64_t search = lk & (~0ull << (63 - nm));
  uint32_t lo = b, hi = e;
  while (hi > lo) {
    const uint32_t mid = (lo + hi) / 2;
    if (keys[mid] < search) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_blevel(Side sd, Temp tp, uint32_t T, int level) {
  __shared__ uint32_t ss[256 / kWave + 1];
  __shared__ uint32_t sNode, sOpen;
  const uint32_t nOpen = sd.ctr[C_LEVEL0 + level];
  const uint32_t start = sd.ctr[C_LSTART0 + level];
  const uint32_t *open = tp.levels + start;
  uint32_t *next = tp.levels + start + nOpen;
  if (blockIdx.x == 0 && threadIdx.x == 0) sd.ctr[C_LSTART0 + level + 1] = start + nOpen;
  for (uint32_t b0 = blockIdx.x * 256; b0 < nOpen; b0 += gridDim.x * 256) {
    const uint32_t i = b0 + threadIdx.x;
    uint32_t node = 0, split = 0;
    uint2 r = make_uint2(0, 0);
    bool doSplit = false, multi = false;
    if (i < nOpen) {
      node = open[i];
      r = tp.ranges[node];
      const uint32_t size = r.y - r.x;
      const uint32_t c0 = (uint32_t)(sd.keys[r.x] >> kCloudShift), c1 = (uint32_t)(sd.keys[r.y - 1] >> kCloudShift);
      multi = c0 != c1;
      if (multi) {
        // the highest differing key bit is a cloud bit: the first key at or
        // above the search key find_split64 would look for is the first point
        // of cloud cs, which the offsets give without a search
        const uint32_t cs = c1 & (~0u << (31 - __clz(c0 ^ c1)));
        split = sd.off[cs];
        doSplit = true;
      } else if (size > T) {
        split = find_split64(sd.keys, r.x, r.y);
        doSplit = true;
      } else {
        sd.nodes[node].oc = pack_oc(r.x, size);
      }
    }
    uint32_t tot;
    const uint32_t ex = block_excl_scan<256>(doSplit ? 1u : 0u, ss, &tot);
    if (threadIdx.x == 0 && tot) {
      sNode = atomicAdd(&sd.ctr[C_NODES], 2 * tot);
      sOpen = atomicAdd(&sd.ctr[C_LEVEL0 + level + 1], 2 * tot);
    }
    __syncthreads();
    if (doSplit) {
      const uint32_t child = sNode + 2 * ex;
      sd.nodes[node].oc = pack_oc(child, 0);
      tp.ranges[child] = make_uint2(r.x, split);
      tp.ranges[child + 1] = make_uint2(split, r.y);
      next[sOpen + 2 * ex] = child;
      next[sOpen + 2 * ex + 1] = child + 1;
      if (multi) {
        const uint64_t cl = sd.keys[r.x] >> kCloudShift, cm = sd.keys[split - 1] >> kCloudShift;
        if (cl == cm) sd.rootOf[cl] = child;
        const uint64_t cr = sd.keys[split] >> kCloudShift, ce = sd.keys[r.y - 1] >> kCloudShift;
        if (cr == ce) sd.rootOf[cr] = child + 1;
      }
    }
    __syncthreads();
  }
}

// bottom-up refit, one launch per level from the deepest (no atomics)
__global__ void k_brefit_level(Side sd, Temp tp, int level) {
  const uint32_t n = sd.ctr[C_LEVEL0 + level];
  const uint32_t *ids = tp.levels + sd.ctr[C_LSTART0 + level];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    Node *nd = &sd.nodes[ids[i]];
    const uint64_t oc = nd->oc;
    const uint32_t cnt = oc_count(oc), o = (uint32_t)oc_offset(oc);
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (cnt) {
      for (uint32_t j = 0; j < cnt; j++) {
        const float4 p = sd.spts[o + j];
        const float x[3] = {p.x, p.y, p.z};
        for (int d = 0; d < 3; d++) {
          lo[d] = fminf(lo[d], x[d]);
          hi[d] = fmaxf(hi[d], x[d]);
        }
      }
    } else {
      const Node l = sd.nodes[o], r = sd.nodes[o + 1];
      for (int d = 0; d < 3; d++) { lo[d] = fminf(l.lo[d], r.lo[d]); hi[d] = fmaxf(l.hi[d], r.hi[d]); }
    }
    for (int d = 0; d < 3; d++) { nd->lo[d] = lo[d]; nd->hi[d] = hi[d]; }
  }
}

// The closest point of one query in the tree at `root`, on a caller's
// traversal stack (a private array, or LDS for the shallow stacks): its
// global index (lowest among equal distances), its sorted slot and distance.
// Leaf prims are read from the sorted points, a contiguous run per leaf.
template <typename StackT>
__device__ __forceinline__ bool fcp_forest_on(StackT &stack, const Side &tgt, uint32_t root, float3 q,
                                              int32_t &best, uint32_t &bestSlot, float &bestD) {
  best = -1;
  bestSlot = kNoSlot;
  bestD = INFINITY;
  cubql_bvh3f bvh;
  bvh.nodes = (cubql_node3f *)tgt.nodes;
  bvh.numNodes = 2;
  bvh.primIDs = tgt.ids;  // only for leaf slot offsets: ids are not read
  bvh.numPrims = tgt.n;
  const float4 *spts = tgt.spts;
  auto test = [&](uint32_t slot) {
    const float4 p = spts[slot];
    const float d = cubql_dev::sqr_dist_pt(p.x, p.y, p.z, q.x, q.y, q.z);
    const int32_t pid = (int32_t)__float_as_uint(p.w);
    if (d < bestD || (d == bestD && best >= 0 && pid < best)) {
      bestD = d;
      best = pid;
      bestSlot = slot;
    }
  };
  // seed the radius with the nearest-child leaf, which the traversal then
  // leaves out
  cubql_dev::NoStats st;
  const uint64_t seedLeaf = cubql_dev::nearest_leaf(bvh, q.x, q.y, q.z, root, st);
  const uint32_t seed = cubql_dev::oc_off(seedLeaf);
  for (uint32_t j = 0; j < cubql_dev::oc_cnt(seedLeaf); j++) test(seed + j);
  auto leaf = [&](const uint32_t *ids, uint32_t cnt) -> float {
    const uint32_t off = (uint32_t)(ids - tgt.ids);
    if (off == seed) return INFINITY;
    for (uint32_t j = 0; j < cnt; j++) test(off + j);
    return bestD;
  };
  return cubql_dev::for_each_leaf_on(stack, bvh, q.x, q.y, q.z, bestD, leaf, root, st);
}
template <int STACK>
__device__ __forceinline__ bool fcp_forest(const Side &tgt, uint32_t root, float3 q, int32_t &best,
                                           uint32_t &bestSlot, float &bestD) {
  cubql_dev::PrivateStack<STACK> stack;
  return fcp_forest_on(stack, tgt, root, q, best, bestSlot, bestD);
}

// queries of side `qs` against the forest of side `ts`, in Morton order of qs
template <int BLOCK, int STACK>
__global__ __launch_bounds__(BLOCK) void k_bquery(Side qs, Side ts, int32_t *nn, float *d2, uint32_t *ovf) {
  // queries sorted by (cloud, Morton key): one contiguous run per XCD
  const uint32_t t = xcd_chunk(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
  if (t >= qs.n) return;
  const float4 qv = qs.spts[t];
  const uint32_t qi = __float_as_uint(qv.w);
  const uint32_t cloud = (uint32_t)(qs.keys[t] >> kCloudShift);
  const uint32_t root = ts.rootOf[cloud];
  int32_t best = -1;
  uint32_t bestSlot = kNoSlot;
  float bestD = INFINITY;
  if (root != kNoRoot) {
    const float3 q = make_float3(qv.x, qv.y, qv.z);
    bool done;
    if constexpr (STACK <= 16) {
      // shallow stacks live in LDS (at most 32 KB a workgroup), deeper
      // ones in the lane's private memory
      __shared__ uint2 sStack[STACK * BLOCK];
      cubql_dev::LdsStack<STACK, BLOCK> stack(sStack + threadIdx.x);
      done = fcp_forest_on(stack, ts, root, q, best, bestSlot, bestD);
    } else {
      done = fcp_forest<STACK>(ts, root, q, best, bestSlot, bestD);
    }
    if (!done) {
      ovf[2 + atomicAdd(&ovf[0], 1u)] = t;
      return;
    }
  }
  nn[qi] = best;
  d2[qi] = bestD;
  qs.nnS[t] = bestSlot;
  qs.d2S[t] = bestD;
}

__global__ __launch_bounds__(64) void k_bquery_deep(Side qs, Side ts, int32_t *nn, float *d2, uint32_t *ovf) {
  const uint32_t n = ovf[0];
  for (uint32_t i = blockIdx.x * 64 + threadIdx.x; i < n; i += gridDim.x * 64) {
    const uint32_t t = ovf[2 + i];
    const float4 qv = qs.spts[t];
    const uint32_t qi = __float_as_uint(qv.w);
    const uint32_t cloud = (uint32_t)(qs.keys[t] >> kCloudShift);
    const float3 q = make_float3(qv.x, qv.y, qv.z);
    int32_t best;
    uint32_t bestSlot;
    float bestD;
    if (!fcp_forest<1024>(ts, ts.rootOf[cloud], q, best, bestSlot, bestD)) {
      // deeper than the deep stack (a cloud's Morton tree is at most 46
      // levels deep, so this does not happen): marked incomplete
      atomicAdd(&ovf[1], 1u);
      best = -2;
      bestSlot = kNoSlot;
      bestD = INFINITY;
    }
    nn[qi] = best;
    d2[qi] = bestD;
    qs.nnS[t] = bestSlot;
    qs.d2S[t] = bestD;
  }
}

// adjoint of loss[c] = mean over the query cloud of d(q, nearest).
// The query's own term goes to gradQ[qi] without an atomic: in the first
// direction (ownFirst) every point of the side is a query exactly once, so
// a store initialises gradQ; in the second only this lane touches
// gradQ[qi] (the other lanes scatter into the other side's array), so a
// plain read-modify-write adds to what the first direction scattered.  The
// neighbour's term is scattered into gradT: queries run in Morton order, so
// lanes next to each other often share their nearest point; a segmented
// wave sum over runs of equal neighbour ids leaves one atomic per run.
// The query, its result and the neighbour are read by sorted position.
__global__ __launch_bounds__(256) void k_bback(Side qs, Side ts, float *loss, float *gradQ, float *gradT,
                                               bool ownFirst) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  const bool act = t < qs.n;
  const uint32_t lane = lane_id();
  uint32_t qi = 0, cloud = 0;
  int32_t p = -1;
  float contrib = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
  if (act) {
    const float4 qv = qs.spts[t];
    qi = __float_as_uint(qv.w);
    cloud = (uint32_t)(qs.keys[t] >> kCloudShift);
    const uint32_t ps = qs.nnS[t];
    if (ps != kNoSlot) {
      const float4 pv = ts.spts[ps];
      p = (int32_t)__float_as_uint(pv.w);
      const float d = sqrtf(qs.d2S[t]);
      const float g = 1.f / (float)(qs.off[cloud + 1] - qs.off[cloud]);
      contrib = g * d;
      if (d > 0.f) {
        const float s = g / d;
        gx = s * (qv.x - pv.x);
        gy = s * (qv.y - pv.y);
        gz = s * (qv.z - pv.z);
      }
    }
    if (gradQ) {
      float *o = gradQ + 3 * (size_t)qi;
      if (ownFirst) {
        o[0] = gx; o[1] = gy; o[2] = gz;
      } else {
        o[0] += gx; o[1] += gy; o[2] += gz;
      }
    }
  }
  if (gradT) {
    // segmented inclusive sums over runs of equal p (a run starts at lane 0
    // or where p differs from the lane before)
    const int32_t pPrev = __shfl_up(p, 1, kWave);
    bool seen = lane == 0 || pPrev != p;
    float sx = gx, sy = gy, sz = gz;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const float ux = __shfl_up(sx, o, kWave), uy = __shfl_up(sy, o, kWave), uz = __shfl_up(sz, o, kWave);
      const bool us = __shfl_up((int)seen, o, kWave) != 0;
      if (lane >= (uint32_t)o && !seen) {
        sx += ux; sy += uy; sz += uz;
        seen = us;
      }
    }
    const int32_t pNext = __shfl_down(p, 1, kWave);
    const bool runEnd = lane == kWave - 1 || pNext != p;
    if (runEnd && p >= 0 && (sx != 0.f || sy != 0.f || sz != 0.f)) {
      atomicAdd(&gradT[3 * (size_t)p + 0], -sx);
      atomicAdd(&gradT[3 * (size_t)p + 1], -sy);
      atomicAdd(&gradT[3 * (size_t)p + 2], -sz);
    }
  }
  if (loss) {
    // one atomic per wave when all active lanes share the cloud
    const uint32_t c0 = __shfl(cloud, 0, kWave);
    const bool uniform = __all(!act || cloud == c0);
    if (uniform) {
      float sum = contrib;
      for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, kWave);
      if (lane == 0 && __any(act)) atomicAdd(&loss[c0], sum);
    } else if (act && contrib != 0.f) {
      atomicAdd(&loss[cloud], contrib);
    }
  }
}

__global__ void k_fill_f32(float *p, float v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ----------------------------------------------------------- workspace
struct Layout {
  size_t bytes = 0;
  size_t take(size_t b) {
    size_t o = bytes;
    bytes += (b + 255) & ~size_t(255);
    return o;
  }
};

struct Plan {
  size_t keys[2], ids[2], spts[2], nnS[2], d2S[2], nodes[2], roots[2], ctr[2];
  size_t keysB, valsB, ranges, levels, sortWs, ovf;
  size_t total;
  uint32_t maxLevels;
};

Plan make_plan(uint32_t numClouds, uint32_t nSrc, uint32_t nTgt) {
  Layout L;
  Plan p{};
  const uint32_t nMax = std::max(std::max(nSrc, nTgt), 1u);
  const uint32_t ns[2] = {nSrc, nTgt};
  for (int s = 0; s < 2; s++) {
    const size_t n = std::max(ns[s], 1u);
    p.keys[s] = L.take(8 * n);
    p.ids[s] = L.take(4 * n);
    p.spts[s] = L.take(16 * n);
    p.nnS[s] = L.take(4 * n);
    p.d2S[s] = L.take(4 * n);
    p.nodes[s] = L.take(32 * (2 * n + 2));
    p.roots[s] = L.take(4 * (size_t)std::max(numClouds, 1u));
    p.ctr[s] = L.take(4 * C_WORDS);
  }
  p.keysB = L.take(8 * (size_t)nMax);
  p.valsB = L.take(4 * (size_t)nMax);
  p.ranges = L.take(8 * (2 * (size_t)nMax + 2));
  p.levels = L.take(4 * (2 * (size_t)nMax + 2));
  int cloudBits = 0;
  while ((1u << cloudBits) < std::max(numClouds, 1u)) cloudBits++;
  p.sortWs = L.take(radix_sort_ws_bytes(nMax, kCloudShift + cloudBits));
  p.ovf = L.take(4 * ((size_t)nMax + 2));
  p.total = L.bytes;
  return p;
}

int cloud_bits(uint32_t numClouds) {
  int b = 0;
  while ((1u << b) < std::max(numClouds, 1u)) b++;
  return b;
}

struct Args {
  const float *src;
  const uint32_t *srcOff;
  uint32_t nSrc;
  const float *tgt;
  const uint32_t *tgtOff;
  uint32_t nTgt;
  uint32_t numClouds;
  cubql_batch_params prm;
  char *ws;
  int32_t *srcNN, *tgtNN;
  float *srcD2, *tgtD2, *loss, *gradSrc, *gradTgt;
};

int build_side(Side &sd, Temp &tp, const Args &a, const Plan &pl, hipStream_t s, int numCU) {
  const int cbits = cloud_bits(a.numClouds);
  CUBQL_CHECK(fill_u32(sd.ctr, 0u, C_WORDS, s));
  CUBQL_CHECK(fill_u32(&sd.ctr[C_N], sd.n, 1, s));
  if (a.numClouds) k_bkeys<<<a.numClouds, 256, 0, s>>>(sd, a.numClouds);
  uint64_t *kr = nullptr;
  uint32_t *vr = nullptr;
  CUBQL_CHECK(radix_sort_pairs_ws_u64(sd.keys, sd.ids, tp.keysB, tp.valsB, &sd.ctr[C_N], std::max(sd.n, 1u),
                                      kCloudShift + cbits, s, tp.sortWs, &kr, &vr));
  if (kr != sd.keys) {
    CUBQL_HIP(hipMemcpyAsync(sd.keys, kr, 8 * (size_t)sd.n, hipMemcpyDeviceToDevice, s));
    CUBQL_HIP(hipMemcpyAsync(sd.ids, vr, 4 * (size_t)sd.n, hipMemcpyDeviceToDevice, s));
  }
  const uint32_t grid = std::max<uint32_t>(1u, std::min<uint32_t>(div_up(sd.n, 256), (uint32_t)numCU * 4));
  if (sd.n) k_bgather<<<grid, 256, 0, s>>>(sd);
  k_binit<<<std::max(1u, std::min<uint32_t>(div_up(a.numClouds, 256), 64)), 256, 0, s>>>(sd, tp, a.numClouds);
  k_bsingle_root<<<1, 64, 0, s>>>(sd);
  // depth bound: cloud bits + 30 Morton bits + mid splits of equal keys
  int midBits = 0;
  const uint32_t maxCloud = a.prm.maxCloudSize ? a.prm.maxCloudSize : std::max(sd.n, 1u);
  while ((1ull << midBits) < maxCloud) midBits++;
  const int levels = std::min(cbits + kCloudShift + midBits + 2, kMaxLevels);
  const uint32_t T = (uint32_t)std::max(1, std::min(a.prm.leafSize, 65535));
  for (int l = 0; l < levels; l++) k_blevel<<<grid, 256, 0, s>>>(sd, tp, T, l);
  for (int l = levels - 1; l >= 0; l--) k_brefit_level<<<grid, 256, 0, s>>>(sd, tp, l);
  CUBQL_LAUNCH_CHECK();
  return CUBQL_OK;
}

template <int BLOCK, int STACK>
void launch_query(const Side &q, const Side &t, int32_t *nn, float *d2, uint32_t *ovf, hipStream_t s) {
  k_bquery<BLOCK, STACK><<<div_up(std::max(q.n, 1u), BLOCK), BLOCK, 0, s>>>(q, t, nn, d2, ovf);
}

template <int STACK>
void launch_query_b(int block, const Side &q, const Side &t, int32_t *nn, float *d2, uint32_t *ovf,
                    hipStream_t s) {
  if (block == 64) launch_query<64, STACK>(q, t, nn, d2, ovf, s);
  else if (block == 128) launch_query<128, STACK>(q, t, nn, d2, ovf, s);
  else launch_query<256, STACK>(q, t, nn, d2, ovf, s);
}

int query_dir(const Side &q, const Side &t, int32_t *nn, float *d2, uint32_t *ovf, const cubql_batch_params &p,
              hipStream_t s) {
  if (q.n == 0) return CUBQL_OK;
  CUBQL_CHECK(fill_u32(ovf, 0u, 2, s));
  const int b = p.blockSize, st = p.stackDepth;
  if (st == 8) launch_query_b<8>(b, q, t, nn, d2, ovf, s);
  else if (st == 16) launch_query_b<16>(b, q, t, nn, d2, ovf, s);
  else if (st == 32) launch_query_b<32>(b, q, t, nn, d2, ovf, s);
  else launch_query_b<64>(b, q, t, nn, d2, ovf, s);
  k_bquery_deep<<<64, 64, 0, s>>>(q, t, nn, d2, ovf);
  CUBQL_LAUNCH_CHECK();
  return CUBQL_OK;
}

int run(const Args &a, hipStream_t s) {
  const Plan pl = make_plan(a.numClouds, a.nSrc, a.nTgt);
  int dev = 0, numCU = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&numCU, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    numCU = 256;
  Side side[2];
  const float *pts[2] = {a.src, a.tgt};
  const uint32_t *offs[2] = {a.srcOff, a.tgtOff};
  const uint32_t ns[2] = {a.nSrc, a.nTgt};
  for (int i = 0; i < 2; i++) {
    side[i].pts = pts[i];
    side[i].off = offs[i];
    side[i].n = ns[i];
    side[i].keys = (uint64_t *)(a.ws + pl.keys[i]);
    side[i].ids = (uint32_t *)(a.ws + pl.ids[i]);
    side[i].spts = (float4 *)(a.ws + pl.spts[i]);
    side[i].nnS = (uint32_t *)(a.ws + pl.nnS[i]);
    side[i].d2S = (float *)(a.ws + pl.d2S[i]);
    side[i].nodes = (Node *)(a.ws + pl.nodes[i]);
    side[i].rootOf = (uint32_t *)(a.ws + pl.roots[i]);
    side[i].ctr = (uint32_t *)(a.ws + pl.ctr[i]);
  }
  Temp tp;
  tp.keysB = (uint64_t *)(a.ws + pl.keysB);
  tp.valsB = (uint32_t *)(a.ws + pl.valsB);
  tp.ranges = (uint2 *)(a.ws + pl.ranges);
  tp.levels = (uint32_t *)(a.ws + pl.levels);
  tp.sortWs = a.ws + pl.sortWs;
  tp.ovf = (uint32_t *)(a.ws + pl.ovf);

  int ps = prof_begin(s);
  CUBQL_CHECK(build_side(side[0], tp, a, pl, s, numCU));
  CUBQL_CHECK(build_side(side[1], tp, a, pl, s, numCU));
  prof_end(ps, s, "batch_build");
  prof_set(ps, (12.0 + 12.0 + 8.0 * 2 * 6 + 8.0 + 24.0 + 32.0) * (a.nSrc + a.nTgt), a.nSrc + a.nTgt);
  const uint32_t fg = 1024;
  if (a.loss) k_fill_f32<<<std::max(1u, std::min(div_up(a.numClouds, 256), fg)), 256, 0, s>>>(a.loss, 0.f, a.numClouds);
  // gradSrc is initialised by the first backward direction's stores; an
  // empty source side leaves it nothing to initialise
  if (a.gradSrc && !a.nSrc) k_fill_f32<<<fg, 256, 0, s>>>(a.gradSrc, 0.f, 3 * (size_t)a.nSrc);
  if (a.gradTgt) k_fill_f32<<<fg, 256, 0, s>>>(a.gradTgt, 0.f, 3 * (size_t)a.nTgt);
  ps = prof_begin(s);
  CUBQL_CHECK(query_dir(side[0], side[1], a.srcNN, a.srcD2, tp.ovf, a.prm, s));
  CUBQL_CHECK(query_dir(side[1], side[0], a.tgtNN, a.tgtD2, tp.ovf, a.prm, s));
  prof_end(ps, s, "batch_query");
  prof_set(ps, 20.0 * (a.nSrc + a.nTgt), a.nSrc + a.nTgt);
  ps = prof_begin(s);
  if (a.loss || a.gradSrc || a.gradTgt) {
    if (a.nSrc)
      k_bback<<<div_up(a.nSrc, 256), 256, 0, s>>>(side[0], side[1], a.loss, a.gradSrc, a.gradTgt, true);
    if (a.nTgt)
      k_bback<<<div_up(a.nTgt, 256), 256, 0, s>>>(side[1], side[0], a.loss, a.gradTgt, a.gradSrc, false);
  }
  prof_end(ps, s, "batch_backward");
  prof_set(ps, 68.0 * (a.nSrc + a.nTgt), a.nSrc + a.nTgt);
  CUBQL_LAUNCH_CHECK();
  return CUBQL_OK;
}

int check_args(const Args &a, size_t wsBytes) {
  if (a.numClouds == 0) return CUBQL_OK;
  if (!a.src || !a.srcOff || !a.tgt || !a.tgtOff || !a.ws || !a.srcNN || !a.tgtNN || !a.srcD2 || !a.tgtD2) {
    set_error("cubql_batch_chamfer_f3: null argument");
    return CUBQL_ERR_INVALID_ARG;
  }
  const Plan pl = make_plan(a.numClouds, a.nSrc, a.nTgt);
  if (wsBytes < pl.total) {
    set_error("cubql_batch_chamfer_f3: workspace of %zu bytes is below the %zu needed", wsBytes, pl.total);
    return CUBQL_ERR_INVALID_ARG;
  }
  const int b = a.prm.blockSize;
  if (!(b == 64 || b == 128 || b == 256) || !(a.prm.stackDepth == 8 || a.prm.stackDepth == 16 || a.prm.stackDepth == 32 || a.prm.stackDepth == 64) ||
      a.prm.leafSize < 1) {
    set_error("cubql_batch_chamfer_f3: params must have blockSize 64/128/256, stackDepth 8/16/32/64, leafSize >= 1");
    return CUBQL_ERR_INVALID_ARG;
  }
  return CUBQL_OK;
}

}  // namespace bt
}  // namespace cubql

using namespace cubql;

struct cubql_graph_s {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" size_t cubql_batch_chamfer_workspace_f3(uint32_t numClouds, uint32_t numSrc, uint32_t numTgt) {
  return bt::make_plan(numClouds, numSrc, numTgt).total;
}

static bt::Args make_args(const float *src, const uint32_t *srcOff, uint32_t nSrc, const float *tgt,
                          const uint32_t *tgtOff, uint32_t nTgt, uint32_t numClouds,
                          const cubql_batch_params *prm, void *ws, int32_t *srcNN, float *srcD2,
                          int32_t *tgtNN, float *tgtD2, float *loss, float *gradSrc, float *gradTgt) {
  bt::Args a{};
  a.src = src; a.srcOff = srcOff; a.nSrc = nSrc;
  a.tgt = tgt; a.tgtOff = tgtOff; a.nTgt = nTgt;
  a.numClouds = numClouds;
  if (prm) a.prm = *prm;
  else a.prm = cubql_batch_params{8, 256, 64, 0};
  a.ws = (char *)ws;
  a.srcNN = srcNN; a.srcD2 = srcD2; a.tgtNN = tgtNN; a.tgtD2 = tgtD2;
  a.loss = loss; a.gradSrc = gradSrc; a.gradTgt = gradTgt;
  return a;
}

extern "C" int cubql_batch_chamfer_f3(const float *d_src, const uint32_t *d_srcOffsets, uint32_t numSrc,
                                      const float *d_tgt, const uint32_t *d_tgtOffsets, uint32_t numTgt,
                                      uint32_t numClouds, const cubql_batch_params *params, void *d_workspace,
                                      size_t workspaceBytes, int32_t *d_srcNN, float *d_srcSqrDist,
                                      int32_t *d_tgtNN, float *d_tgtSqrDist, float *d_loss,
                                      float *d_gradSrc, float *d_gradTgt, cubql_stream_t s) {
  clear_error();
  bt::Args a = make_args(d_src, d_srcOffsets, numSrc, d_tgt, d_tgtOffsets, numTgt, numClouds, params,
                         d_workspace, d_srcNN, d_srcSqrDist, d_tgtNN, d_tgtSqrDist, d_loss, d_gradSrc, d_gradTgt);
  CUBQL_CHECK(bt::check_args(a, workspaceBytes));
  if (numClouds == 0) return CUBQL_OK;
  return bt::run(a, (hipStream_t)s);
}

extern "C" int cubql_batch_chamfer_graph_f3(const float *d_src, const uint32_t *d_srcOffsets, uint32_t numSrc,
                                            const float *d_tgt, const uint32_t *d_tgtOffsets, uint32_t numTgt,
                                            uint32_t numClouds, const cubql_batch_params *params,
                                            void *d_workspace, size_t workspaceBytes, int32_t *d_srcNN,
                                            float *d_srcSqrDist, int32_t *d_tgtNN, float *d_tgtSqrDist,
                                            float *d_loss, float *d_gradSrc, float *d_gradTgt,
                                            cubql_graph *out) {
  clear_error();
  if (!out) {
    set_error("cubql_batch_chamfer_graph_f3: out is NULL");
    return CUBQL_ERR_INVALID_ARG;
  }
  *out = nullptr;
  bt::Args a = make_args(d_src, d_srcOffsets, numSrc, d_tgt, d_tgtOffsets, numTgt, numClouds, params,
                         d_workspace, d_srcNN, d_srcSqrDist, d_tgtNN, d_tgtSqrDist, d_loss, d_gradSrc, d_gradTgt);
  CUBQL_CHECK(bt::check_args(a, workspaceBytes));
  if (numClouds == 0) {
    set_error("cubql_batch_chamfer_graph_f3: empty batch");
    return CUBQL_ERR_INVALID_ARG;
  }
  hipStream_t cs;
  CUBQL_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  cubql_graph g = new cubql_graph_s();
  const bool profWas = prof_on();
  if (profWas) cubql_profile_enable(0);
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  int st = e == hipSuccess ? bt::run(a, cs) : CUBQL_ERR_HIP;
  hipError_t e2 = hipStreamEndCapture(cs, &g->graph);
  if (profWas) cubql_profile_enable(1);
  if (st == CUBQL_OK && (e != hipSuccess || e2 != hipSuccess)) {
    set_error("graph capture failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    st = CUBQL_ERR_HIP;
  }
  if (st == CUBQL_OK) {
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      set_error("hipGraphInstantiate failed: %s", hipGetErrorString(e));
      st = CUBQL_ERR_HIP;
    }
  }
  (void)hipStreamDestroy(cs);
  if (st != CUBQL_OK) {
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return st;
  }
  *out = g;
  return CUBQL_OK;
}

extern "C" int cubql_graph_launch(cubql_graph g, cubql_stream_t s) {
  clear_error();
  if (!g || !g->exec) {
    set_error("cubql_graph_launch: invalid graph");
    return CUBQL_ERR_INVALID_ARG;
  }
  CUBQL_HIP(hipGraphLaunch(g->exec, (hipStream_t)s));
  return CUBQL_OK;
}

extern "C" int cubql_graph_destroy(cubql_graph g) {
  if (!g) return CUBQL_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return CUBQL_OK;
}
