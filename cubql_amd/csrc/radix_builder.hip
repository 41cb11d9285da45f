// radix_builder.hip -- Morton-code (radix) BVH builder for gfx950.
//
// Replaces cuBQL::cuda::radixBuilder (builder/cuda/radix.h:591-771):
//   1. centroid bounds of the valid prims (radix.h:199-241) and the uniform
//      quantizer scale = 2^b * (1/max(largest extent, 1e-20)) (radix.h:41-53);
//   2. Morton key per valid prim: cell = min(u32((c-lo)*scale), 2^b-1),
//      bit i of x,y,z -> bits 3i,3i+1,3i+2 (radix.h:54-62, 283-394);
//      valid prims are compacted in ascending primID order (the reference's
//      atomic tail swap, radix.h:412-416, is order-nondeterministic);
//   3. stable key/value radix sort (radix_sort.hip, replacing cub SortPairs);
//   4. top-down hierarchy: a range of <= T prims is a leaf, a range of equal
//      keys is split in the middle, otherwise at the first key that has the
//      highest differing bit of its first and last key set (findSplit,
//      radix.h:446-482; createNodes radix.h:501-570), one launch per level,
//      children allocated with one atomic per workgroup;
//   5. exact bottom-up refit (refit.hip).
// 63-bit keys (21 bits per axis) are the reference default; 30-bit keys
// (10 bits per axis) are its radix.h:34-38 variant.
#include "block_ops.h"
#include "build_common.h"
#include "wave_ops.h"
#include "common.h"
#include "profile.h"
#include "radix_sort.h"
#include "refit.h"

#include <vector>

namespace cubql {
namespace rb {

constexpr int kBlock = 256;
constexpr int kItems = 8;
constexpr int kTile = kBlock * kItems;
constexpr int kLevelsPerSync = 8;
// > 63 key bits + log2(2^31) mid splits; the rebin builder restarts the
// level sequence after each rebin round, so it gets more room
constexpr int kMaxLevels = 1024;

// counter slots: [C_LEVEL0 + l] = node count of level l, [C_LSTART0 + l] =
// start of level l in the all-levels node list; C_JOBS.. are used by the
// rebin builder only
enum {
  C_NUMVALID = 0, C_NODECOUNT = 1, C_JOBS = 2, C_JOBPRIMS = 3, C_SMALL = 4, C_CURJOBS = 5,
  C_NODESOUT = 6,  // node count after the thread phase
  C_LEVEL0 = 8, C_LSTART0 = C_LEVEL0 + kMaxLevels + 4
};

struct RCtx {
  const cubql_box3f *boxes;
  uint32_t n, T;
  int bitsPerDim;
  uint32_t *blockCount;
  float *blockBounds;
  uint32_t *ctr;      // [C_LEVEL0 + level]: open-node count of each level
  float *quant;       // lo[3], scale
  uint32_t *ids;      // compacted primIDs (sorted in place by the sort)
  Node *nodes;        // temp nodes, capacity 2*n+2
  uint2 *ranges;      // [b,e) per node
  uint32_t *levels;   // node ids level by level (capacity 2n+2, more with rebinning)
  // rebin builder
  uint32_t *jobs;      // nodes whose keys are all equal (> 32 prims)
  uint32_t *jobStart;  // exclusive prefix of job sizes
  int32_t *jobDom;     // per job: encoded centroid bounds
  float *jobQuant;     // per job: lo xyz, scale
  uint32_t *jobLocal;  // per job: 1 = key by local index
  uint32_t *small;     // thread-phase nodes (<= 32 prims)
  uint32_t *smallCount, *smallBase;
  Node *loc;           // thread-phase local nodes at loc[2*begin + j]
  uint64_t *scanStatus;  // look-back status of the thread-phase id scan
};

__global__ __launch_bounds__(kBlock) void k_rcount(RCtx c) {
  __shared__ float sb[12 * (kBlock / kWave)];
  __shared__ uint32_t ss[kBlock / kWave + 1];
  const uint32_t base = blockIdx.x * kTile;
  uint32_t cnt = 0;
  float v[12];
  for (int j = 0; j < 3; j++) { v[j] = INFINITY; v[j + 3] = -INFINITY; v[j + 6] = INFINITY; v[j + 9] = -INFINITY; }
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    uint32_t i = base + k * kBlock + threadIdx.x;
    if (i < c.n) {
      cubql_box3f b = load_box(c.boxes, i);
      if (box_is_valid(b)) {
        cnt++;
        for (int d = 0; d < 3; d++) {
          float ce = 0.5f * (b.lower[d] + b.upper[d]);
          v[d] = fminf(v[d], ce);
          v[d + 3] = fmaxf(v[d + 3], ce);
        }
      }
    }
  }
  block_reduce_bounds<kBlock>(v, sb);
  uint32_t tot;
  block_excl_scan<kBlock>(cnt, ss, &tot);
  if (threadIdx.x < 6) c.blockBounds[blockIdx.x * 8 + threadIdx.x] = v[threadIdx.x];
  if (threadIdx.x == 0) c.blockCount[blockIdx.x] = tot;
}

// single workgroup: tile offsets, centroid bounds, quantizer, root node
__global__ __launch_bounds__(1024) void k_rscan(RCtx c, uint32_t numBlocks) {
  __shared__ uint32_t ss[1024 / kWave + 1];
  __shared__ float sr[6][1024 / kWave];
  uint32_t run = 0;
  float red[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (uint32_t b0 = 0; b0 < numBlocks; b0 += 1024) {
    uint32_t i = b0 + threadIdx.x;
    uint32_t v = i < numBlocks ? c.blockCount[i] : 0;
    if (i < numBlocks) {
      const float *bb = c.blockBounds + (size_t)i * 8;
      for (int j = 0; j < 3; j++) { red[j] = fminf(red[j], bb[j]); red[j + 3] = fmaxf(red[j + 3], bb[j + 3]); }
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan<1024>(v, ss, &tot);
    if (i < numBlocks) c.blockCount[i] = run + ex;
    run += tot;
  }
  for (int j = 0; j < 6; j++) {
    float r = j < 3 ? wave_min(red[j]) : wave_max(red[j]);
    if (lane_id() == 0) sr[j][threadIdx.x / kWave] = r;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float lo[3], hi[3];
    for (int j = 0; j < 6; j++) {
      float r = sr[j][0];
      for (int w = 1; w < 1024 / kWave; w++) r = j < 3 ? fminf(r, sr[j][w]) : fmaxf(r, sr[j][w]);
      (j < 3 ? lo[j] : hi[j - 3]) = r;
    }
    // Quantizer::init (radix.h:46-53)
    float ext = fmaxf(fmaxf(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
    float inv = 1.f / fmaxf(ext, 1e-20f);
    float scale = (float)(1u << c.bitsPerDim) * inv;
    c.quant[0] = lo[0]; c.quant[1] = lo[1]; c.quant[2] = lo[2]; c.quant[3] = scale;
    c.ctr[C_NUMVALID] = run;
    c.ctr[C_NODECOUNT] = 2;
    Node z;
    for (int d = 0; d < 3; d++) { z.lo[d] = 0.f; z.hi[d] = 0.f; }
    z.oc = 0;
    c.nodes[1] = z;  // radix.h:585-586
    if (run > 0) {
      c.ranges[0] = make_uint2(0u, run);
      c.levels[0] = 0;
      c.ctr[C_LEVEL0] = 1;
    } else {
      c.ctr[C_LEVEL0] = 0;
    }
  }
}

__device__ __forceinline__ uint64_t spread21(uint32_t v) {
  uint64_t x = v & 0x1fffffu;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  uint32_t x = v & 0x3ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

template <typename K>
__device__ __forceinline__ K morton(const float c[3], const float *q, int bits) {
  const uint32_t cmax = (1u << bits) - 1;
  uint32_t cell[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    float f = (c[d] - q[d]) * q[3];  // radix.h:59, (P-bias)*scale
    uint32_t u = (uint32_t)f;
    cell[d] = min(u, cmax);
  }
  if (sizeof(K) == 8)
    return (K)(spread21(cell[0]) | (spread21(cell[1]) << 1) | (spread21(cell[2]) << 2));
  return (K)(spread10(cell[0]) | (spread10(cell[1]) << 1) | (spread10(cell[2]) << 2));
}

template <typename K>
__global__ __launch_bounds__(kBlock) void k_rkeys(RCtx c, K *keys) {
  __shared__ uint32_t sr[kItems * (kBlock / kWave) + 1];
  __shared__ float sq[4];
  if (threadIdx.x < 4) sq[threadIdx.x] = c.quant[threadIdx.x];
  __syncthreads();
  const uint32_t base = blockIdx.x * kTile;
  bool valid[kItems];
  K key[kItems];
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    uint32_t i = base + k * kBlock + threadIdx.x;
    valid[k] = false;
    key[k] = 0;
    if (i < c.n) {
      cubql_box3f b = load_box(c.boxes, i);
      valid[k] = box_is_valid(b);
      float ce[3] = {0.5f * (b.lower[0] + b.upper[0]), 0.5f * (b.lower[1] + b.upper[1]),
                     0.5f * (b.lower[2] + b.upper[2])};
      key[k] = morton<K>(ce, sq, c.bitsPerDim);
    }
  }
  uint32_t rank[kItems];
  block_rank_striped<kBlock, kItems>(valid, rank, sr);
  const uint32_t off = c.blockCount[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    if (!valid[k]) continue;
    const uint32_t i = base + k * kBlock + threadIdx.x;
    keys[off + rank[k]] = key[k];
    c.ids[off + rank[k]] = i;
  }
}

template <typename K>
__device__ __forceinline__ uint32_t find_split(const K *keys, uint32_t b, uint32_t e) {
  const K fk = keys[b], lk = keys[e - 1];
  if (fk == lk) return (uint32_t)(((uint64_t)b + e) / 2);  // radix.h:455-465
  const int nbits = 8 * (int)sizeof(K);
  const int nm = sizeof(K) == 8 ? __clzll((unsigned long long)(fk ^ lk)) : __clz((uint32_t)(fk ^ lk));
  const K search = lk & (~(K)0 << (nbits - 1 - nm));
  uint32_t lo = b, hi = e;
  while (hi > lo) {
    const uint32_t mid = (lo + hi) / 2;
    if (keys[mid] < search) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// One tree level: each open node becomes a leaf or splits in two.
template <typename K>
__global__ __launch_bounds__(kBlock) void k_rlevel(RCtx c, const K *keys, int level) {
  __shared__ uint32_t ss[kBlock / kWave + 1];
  __shared__ uint32_t sNode, sOpen;
  const uint32_t nOpen = c.ctr[C_LEVEL0 + level];
  const uint32_t start = c.ctr[C_LSTART0 + level];
  const uint32_t *open = c.levels + start;
  uint32_t *next = c.levels + start + nOpen;  // the next level follows this one
  if (blockIdx.x == 0 && threadIdx.x == 0) c.ctr[C_LSTART0 + level + 1] = start + nOpen;
  for (uint32_t b0 = blockIdx.x * kBlock; b0 < nOpen; b0 += gridDim.x * kBlock) {
    const uint32_t i = b0 + threadIdx.x;
    uint32_t node = 0, split = 0;
    uint2 r = make_uint2(0, 0);
    bool doSplit = false;
    if (i < nOpen) {
      node = open[i];
      r = c.ranges[node];
      const uint32_t size = r.y - r.x;
      if (size > c.T) {
        split = find_split<K>(keys, r.x, r.y);
        doSplit = true;
      } else {
        c.nodes[node].oc = pack_oc(r.x, size);
      }
    }
    uint32_t tot;
    const uint32_t ex = block_excl_scan<kBlock>(doSplit ? 1u : 0u, ss, &tot);
    if (threadIdx.x == 0 && tot) {
      sNode = atomicAdd(&c.ctr[C_NODECOUNT], 2 * tot);
      sOpen = atomicAdd(&c.ctr[C_LEVEL0 + level + 1], 2 * tot);
    }
    __syncthreads();
    if (doSplit) {
      const uint32_t child = sNode + 2 * ex;
      c.nodes[node].oc = pack_oc(child, 0);
      c.ranges[child] = make_uint2(r.x, split);
      c.ranges[child + 1] = make_uint2(split, r.y);
      next[sOpen + 2 * ex] = child;
      next[sOpen + 2 * ex + 1] = child + 1;
    }
    __syncthreads();
  }
}

__global__ void k_rinit_levels(uint32_t *ctr) {
  for (int i = threadIdx.x; i < kMaxLevels + 2; i += blockDim.x) ctr[C_LEVEL0 + 1 + i] = 0;
  if (threadIdx.x < C_LEVEL0 - C_JOBS) ctr[C_JOBS + threadIdx.x] = 0;
  if (threadIdx.x == 0) ctr[C_LSTART0] = 0;
}

// Bottom-up refit from the deepest level: a leaf takes the union of its prim
// boxes, an inner node the union of its children (refitted before it).
__device__ __forceinline__ void refit_node(const RCtx &c, Node *nodes, const uint32_t *primIDs, uint32_t id) {
  Node *nd = &nodes[id];
  const uint64_t oc = nd->oc;
  const uint32_t cnt = oc_count(oc), off = (uint32_t)oc_offset(oc);
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  if (cnt) {
    for (uint32_t j = 0; j < cnt; j++) {
      const cubql_box3f b = load_box(c.boxes, primIDs[off + j]);
      for (int d = 0; d < 3; d++) { lo[d] = fminf(lo[d], b.lower[d]); hi[d] = fmaxf(hi[d], b.upper[d]); }
    }
  } else {
    const Node l = nodes[off], r = nodes[off + 1];
    for (int d = 0; d < 3; d++) { lo[d] = fminf(l.lo[d], r.lo[d]); hi[d] = fmaxf(l.hi[d], r.hi[d]); }
  }
  for (int d = 0; d < 3; d++) { nd->lo[d] = lo[d]; nd->hi[d] = hi[d]; }
}

// one launch per wide level (the launch boundary orders the levels)
__global__ void k_rrefit_level(RCtx c, Node *nodes, const uint32_t *primIDs, int level) {
  const uint32_t n = c.ctr[C_LEVEL0 + level];
  const uint32_t *ids = c.levels + c.ctr[C_LSTART0 + level];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    refit_node(c, nodes, primIDs, ids[i]);
}

// the narrow top levels [0, levels) in one workgroup, deepest first; its
// barriers order the levels (one CU: the node writes are seen by the
// workgroup's later reads)
constexpr uint32_t kRefitTopMax = 4096;  // nodes per level for the one-workgroup pass
__global__ __launch_bounds__(1024) void k_rrefit_top(RCtx c, Node *nodes, const uint32_t *primIDs, int levels) {
  for (int level = levels - 1; level >= 0; level--) {
    const uint32_t n = c.ctr[C_LEVEL0 + level];
    const uint32_t *ids = c.levels + c.ctr[C_LSTART0 + level];
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) refit_node(c, nodes, primIDs, ids[i]);
    __syncthreads();
  }
}

// host: refit levels [0, levels) given their node counts
static void refit_launch(const RCtx &c, Node *nodes, const uint32_t *primIDs, int levels,
                         const std::vector<uint32_t> &hc, int numCU, hipStream_t s) {
  int top = 0;  // levels [0, top) are narrow
  while (top < levels && hc[C_LEVEL0 + top] <= kRefitTopMax) top++;
  for (int l = levels - 1; l >= top; l--) {
    const uint32_t g = std::max<uint32_t>(1u, std::min<uint32_t>(div_up(hc[C_LEVEL0 + l], kBlock), (uint32_t)numCU * 8));
    k_rrefit_level<<<g, kBlock, 0, s>>>(c, nodes, primIDs, l);
  }
  if (top > 0) k_rrefit_top<<<1, 1024, 0, s>>>(c, nodes, primIDs, top);
}

template <typename K>
int build(cubql_bvh3f *bvh, const cubql_box3f *boxes, uint32_t n, uint32_t T, int bitsPerDim,
          hipStream_t s, const cubql_mem_resource *mem) {
  ScratchArena scr(mem, s);
  RCtx c{};
  c.boxes = boxes;
  c.n = n;
  c.T = T;
  c.bitsPerDim = bitsPerDim;
  const uint32_t numBlocks = div_up(n, kTile);
  K *keysA = nullptr, *keysB = nullptr;
  uint32_t *idsB = nullptr;
  CUBQL_CHECK(scr.reserve(&c.blockCount, numBlocks + 1));
  CUBQL_CHECK(scr.reserve(&c.blockBounds, 8 * (size_t)numBlocks + 8));
  CUBQL_CHECK(scr.reserve(&c.ctr, C_LSTART0 + kMaxLevels + 8));
  CUBQL_CHECK(scr.reserve(&c.quant, 8));
  CUBQL_CHECK(scr.reserve(&keysA, n));
  CUBQL_CHECK(scr.reserve(&keysB, n));
  CUBQL_CHECK(scr.reserve(&idsB, n));
  CUBQL_CHECK(scr.reserve(&c.nodes, 2 * (size_t)n + 2));
  CUBQL_CHECK(scr.reserve(&c.ranges, 2 * (size_t)n + 2));
  CUBQL_CHECK(scr.reserve(&c.levels, 2 * (size_t)n + 2));
  CUBQL_CHECK(scr.commit());
  uint32_t *primIDs = nullptr;
  CUBQL_CHECK(mem_alloc(mem, (void **)&primIDs, (size_t)n * sizeof(uint32_t), s));
  c.ids = primIDs;
  auto fail = [&](int st) {
    mem_free(mem, primIDs, s);
    return st;
  };
  k_rinit_levels<<<1, 256, 0, s>>>(c.ctr);
  int ks = prof_begin(s);
  k_rcount<<<numBlocks, kBlock, 0, s>>>(c);
  k_rscan<<<1, 1024, 0, s>>>(c, numBlocks);
  k_rkeys<K><<<numBlocks, kBlock, 0, s>>>(c, keysA);
  prof_end(ks, s, "radix_keys");
  prof_set(ks, 2.0 * 24.0 * n + (sizeof(K) + 4.0) * n, n);
  if (hipGetLastError() != hipSuccess) return fail((set_error("radix key launch failed"), CUBQL_ERR_HIP));
  const int keyBits = 3 * bitsPerDim;
  K *kRes = nullptr;
  uint32_t *vRes = nullptr;
  int st;
  if (sizeof(K) == 8)
    st = radix_sort_pairs_dev_u64((uint64_t *)keysA, primIDs, (uint64_t *)keysB, idsB, &c.ctr[C_NUMVALID],
                                  n, keyBits, s, mem, (uint64_t **)&kRes, &vRes);
  else
    st = radix_sort_pairs_dev_u32((uint32_t *)keysA, primIDs, (uint32_t *)keysB, idsB, &c.ctr[C_NUMVALID],
                                  n, keyBits, s, mem, (uint32_t **)&kRes, &vRes);
  if (st) return fail(st);
  if (vRes != primIDs) {
    if (hipMemcpyAsync(primIDs, vRes, (size_t)n * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
      return fail((set_error("id copy failed"), CUBQL_ERR_HIP));
  }
  // hierarchy, level by level, with a completion check every few levels
  int dev = 0, numCU = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&numCU, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    numCU = 256;
  const uint32_t grid = std::max<uint32_t>(1u, std::min<uint32_t>(div_up(n, kBlock), (uint32_t)numCU * 8));
  std::vector<uint32_t> hc(C_LSTART0 + kMaxLevels + 8);
  int level = 0;
  int hs = prof_begin(s);
  while (true) {
    for (int i = 0; i < kLevelsPerSync && level < kMaxLevels; i++, level++) {
      // level l has at most 2^l open nodes: the top levels launch a few
      // workgroups instead of a chip-wide grid of idle ones
      const uint64_t maxOpen = std::min<uint64_t>(level < 40 ? (1ull << level) : ~0ull, (uint64_t)n);
      const uint32_t gridL = std::max<uint32_t>(1u, std::min<uint32_t>(grid, div_up(maxOpen, kBlock)));
      k_rlevel<K><<<gridL, kBlock, 0, s>>>(c, kRes, level);
    }
    if (hipGetLastError() != hipSuccess) return fail((set_error("level launch failed"), CUBQL_ERR_HIP));
    if (hipMemcpyAsync(hc.data(), c.ctr, sizeof(uint32_t) * hc.size(), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return fail((set_error("level sync failed"), CUBQL_ERR_HIP));
    if (hc[C_LEVEL0 + level] == 0) break;
    if (level >= kMaxLevels) return fail((set_error("radix hierarchy deeper than %d levels", kMaxLevels), CUBQL_ERR_INTERNAL));
  }
  prof_end(hs, s, "radix_hierarchy");
  const uint32_t numValid = hc[C_NUMVALID];
  const uint32_t numNodes = numValid ? hc[C_NODECOUNT] : 0;
  prof_set(hs, 8.0 * numNodes, numNodes);
  Node *nodes = nullptr;
  st = mem_alloc(mem, (void **)&nodes, std::max<size_t>(numNodes, 2) * sizeof(Node), s);
  if (st) return fail(st);
  if (numNodes &&
      hipMemcpyAsync(nodes, c.nodes, (size_t)numNodes * sizeof(Node), hipMemcpyDeviceToDevice, s) != hipSuccess) {
    mem_free(mem, nodes, s);
    return fail((set_error("node copy failed"), CUBQL_ERR_HIP));
  }
  bvh->nodes = (cubql_node3f *)nodes;
  bvh->numNodes = numNodes;
  bvh->primIDs = primIDs;
  bvh->numPrims = numValid;
  int rs = prof_begin(s);
  refit_launch(c, nodes, primIDs, level, hc, numCU, s);
  st = hipGetLastError() == hipSuccess ? CUBQL_OK : (set_error("refit launch failed"), CUBQL_ERR_HIP);
  prof_end(rs, s, "refit");
  prof_set(rs, 24.0 * numValid + 64.0 * numNodes, numNodes);
  if (prof_on()) {
    if (hipStreamSynchronize(s) == hipSuccess) prof_flush();
  }
  return st;
}


// ------------------------------------------------------------------ rebin
// Robust rebinning builder (rebinMortonBuilder.h:595-1476, NO_BLOCK_PHASE
// as in the reference, :22).
constexpr uint32_t kThreadPhase = 32;  // smallEnoughToLeaveBroadPhase (:624-636)

// One broad-phase level: each open node becomes a leaf (size <= T), a
// thread-phase node (size <= 32), splits at findSplit, or -- all keys
// equal -- a rebin job (broadPhaseCreateNodes, :664-749).
__global__ __launch_bounds__(kBlock) void k_rlevel_rebin(RCtx c, const uint32_t *keys, int level) {
  __shared__ uint32_t ss[kBlock / kWave + 1];
  __shared__ uint32_t sNode, sOpen, sSmall, sJob;
  const uint32_t nOpen = c.ctr[C_LEVEL0 + level];
  const uint32_t start = c.ctr[C_LSTART0 + level];
  const uint32_t *open = c.levels + start;
  uint32_t *next = c.levels + start + nOpen;
  if (blockIdx.x == 0 && threadIdx.x == 0) c.ctr[C_LSTART0 + level + 1] = start + nOpen;
  for (uint32_t b0 = blockIdx.x * kBlock; b0 < nOpen; b0 += gridDim.x * kBlock) {
    const uint32_t i = b0 + threadIdx.x;
    uint32_t node = 0, split = 0;
    uint2 r = make_uint2(0, 0);
    int cat = -1;  // 0 split, 1 thread phase, 2 rebin job
    if (i < nOpen) {
      node = open[i];
      r = c.ranges[node];
      const uint32_t size = r.y - r.x;
      if (size <= c.T) {
        c.nodes[node].oc = pack_oc(r.x, size);
      } else if (size <= kThreadPhase) {
        cat = 1;
      } else if (keys[r.x] != keys[r.y - 1]) {
        split = find_split<uint32_t>(keys, r.x, r.y);
        cat = 0;
      } else {
        cat = 2;
      }
    }
    uint32_t totSp, totSm, totJ;
    const uint32_t exSp = block_excl_scan<kBlock>(cat == 0 ? 1u : 0u, ss, &totSp);
    const uint32_t exSm = block_excl_scan<kBlock>(cat == 1 ? 1u : 0u, ss, &totSm);
    const uint32_t exJ = block_excl_scan<kBlock>(cat == 2 ? 1u : 0u, ss, &totJ);
    if (threadIdx.x == 0) {
      if (totSp) {
        sNode = atomicAdd(&c.ctr[C_NODECOUNT], 2 * totSp);
        sOpen = atomicAdd(&c.ctr[C_LEVEL0 + level + 1], 2 * totSp);
      }
      if (totSm) sSmall = atomicAdd(&c.ctr[C_SMALL], totSm);
      if (totJ) sJob = atomicAdd(&c.ctr[C_JOBS], totJ);
    }
    __syncthreads();
    if (cat == 0) {
      const uint32_t child = sNode + 2 * exSp;
      c.nodes[node].oc = pack_oc(child, 0);
      c.ranges[child] = make_uint2(r.x, split);
      c.ranges[child + 1] = make_uint2(split, r.y);
      next[sOpen + 2 * exSp] = child;
      next[sOpen + 2 * exSp + 1] = child + 1;
    } else if (cat == 1) {
      c.small[sSmall + exSm] = node;
    } else if (cat == 2) {
      c.jobs[sJob + exJ] = node;
    }
    __syncthreads();
  }
}

// Single workgroup: job prefix sums; the jobs become the open nodes of
// `level` (the level loop resumes there once their keys are rebuilt).
__global__ __launch_bounds__(1024) void k_rebin_setup(RCtx c, int level) {
  __shared__ uint32_t ss[1024 / kWave + 1];
  const uint32_t J = c.ctr[C_JOBS];
  const uint32_t start = c.ctr[C_LSTART0 + level];
  uint32_t run = 0;
  for (uint32_t b0 = 0; b0 < J; b0 += 1024) {
    const uint32_t j = b0 + threadIdx.x;
    uint32_t size = 0;
    if (j < J) {
      const uint2 r = c.ranges[c.jobs[j]];
      size = r.y - r.x;
      c.levels[start + j] = c.jobs[j];
      for (int d = 0; d < 3; d++) {
        c.jobDom[6 * (size_t)j + d] = enc_f(INFINITY);
        c.jobDom[6 * (size_t)j + 3 + d] = enc_f(-INFINITY);
      }
    }
    uint32_t tot;
    const uint32_t ex = block_excl_scan<1024>(size, ss, &tot);
    if (j < J) c.jobStart[j] = run + ex;
    run += tot;
  }
  if (threadIdx.x == 0) {
    c.jobStart[J] = run;
    c.ctr[C_JOBPRIMS] = run;
    c.ctr[C_CURJOBS] = J;
    c.ctr[C_JOBS] = 0;
    c.ctr[C_LEVEL0 + level] = J;
  }
}

__device__ __forceinline__ void center3(const cubql_box3f &b, float (&ce)[3]) {
  for (int d = 0; d < 3; d++) ce[d] = 0.5f * (b.lower[d] + b.upper[d]);
}

// Quantizer<float,3>::init over a centroid domain (:38-60) into q[4].
__device__ __forceinline__ void quant_init(const float (&lo)[3], const float (&hi)[3], float *q) {
  const float ext = fmaxf(fmaxf(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
  q[0] = lo[0]; q[1] = lo[1]; q[2] = lo[2];
  q[3] = 1024.f * (1.f / fmaxf(ext, 1e-20f));
}

__device__ __forceinline__ uint32_t find_job(const uint32_t *start, uint32_t J, uint32_t t) {
  uint32_t lo = 0, hi = J;  // last j with start[j] <= t
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (start[mid] <= t) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Centroid domain of every job, over all rebin prims at once: lanes of a
// wave whose prims all belong to one job (the common case: jobs are
// contiguous and large) reduce in registers and issue one atomic per bound;
// mixed waves fall back to per-lane atomics.  jobDom holds encoded bounds.
__global__ __launch_bounds__(kBlock) void k_rebin_grow(RCtx c) {
  const uint32_t J = c.ctr[C_CURJOBS];
  const uint32_t P = c.ctr[C_JOBPRIMS];
  const uint32_t lane = lane_id();
  const uint32_t stride = gridDim.x * kBlock;
  for (uint32_t base = blockIdx.x * kBlock + (threadIdx.x & ~(uint32_t)(kWave - 1)); base < P; base += stride) {
    const uint32_t t = base + lane;
    const bool active = t < P;
    uint32_t j = 0;
    float ce[3] = {0.f, 0.f, 0.f};
    if (active) {
      j = find_job(c.jobStart, J, t);
      center3(load_box(c.boxes, c.ids[c.ranges[c.jobs[j]].x + (t - c.jobStart[j])]), ce);
    }
    const uint32_t j0 = __builtin_amdgcn_readfirstlane(j);  // lane 0 is active when base < P
    int32_t *dom = c.jobDom + 6 * (size_t)j;
    if (__ballot(active && j != j0) == 0) {
      float v[6];
      for (int d = 0; d < 3; d++) {
        v[d] = wave_min_dpp(active ? ce[d] : INFINITY);
        v[d + 3] = wave_max_dpp(active ? ce[d] : -INFINITY);
      }
      if (lane < 6) {
        const float x = lane == 0 ? v[0] : lane == 1 ? v[1] : lane == 2 ? v[2] : lane == 3 ? v[3] : lane == 4 ? v[4] : v[5];
        if (lane < 3) atomicMin(&c.jobDom[6 * (size_t)j0 + lane], enc_f(x));
        else atomicMax(&c.jobDom[6 * (size_t)j0 + lane], enc_f(x));
      }
    } else if (active) {
      for (int d = 0; d < 3; d++) {
        atomicMin(&dom[d], enc_f(ce[d]));
        atomicMax(&dom[d + 3], enc_f(ce[d]));
      }
    }
  }
}

// Per job: quantizer of its domain and whether to key by local index: the
// domain is one point (rebinCreateNewKeys, :1052-1055), or all its prims fall
// in one cell (the cells of the domain's lower and upper corner agree on
// every axis), which the reference would rebin forever.
__global__ void k_rebin_finalize(RCtx c) {
  const uint32_t J = c.ctr[C_CURJOBS];
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < J; j += gridDim.x * blockDim.x) {
    float lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
      lo[d] = dec_f(c.jobDom[6 * (size_t)j + d]);
      hi[d] = dec_f(c.jobDom[6 * (size_t)j + 3 + d]);
    }
    float q[4];
    quant_init(lo, hi, q);
    const bool point = lo[0] == hi[0] && lo[1] == hi[1] && lo[2] == hi[2];
    const bool oneCell = morton<uint32_t>(lo, q, 10) == morton<uint32_t>(hi, q, 10);
    for (int k = 0; k < 4; k++) c.jobQuant[4 * j + k] = q[k];
    c.jobLocal[j] = (point || oneCell) ? 1u : 0u;
  }
}

// New (job:key) pairs, one per prim of every job, in position order.
__global__ __launch_bounds__(kBlock) void k_rebin_keys(RCtx c, uint64_t *rk, uint32_t *rv) {
  const uint32_t J = c.ctr[C_CURJOBS];
  const uint32_t P = c.ctr[C_JOBPRIMS];
  for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < P; t += gridDim.x * kBlock) {
    const uint32_t j = find_job(c.jobStart, J, t);
    const uint32_t local = t - c.jobStart[j];
    const uint32_t id = c.ids[c.ranges[c.jobs[j]].x + local];
    uint32_t key = local;
    if (!c.jobLocal[j]) {
      float ce[3];
      center3(load_box(c.boxes, id), ce);
      key = morton<uint32_t>(ce, c.jobQuant + 4 * j, 10);
    }
    rk[t] = ((uint64_t)j << 32) | key;
    rv[t] = id;
  }
}

// Sorted pairs back into each job's range (rebinWriteBack, :1002-1023).
__global__ __launch_bounds__(kBlock) void k_rebin_writeback(RCtx c, uint32_t *keys, const uint64_t *rk,
                                                            const uint32_t *rv) {
  const uint32_t P = c.ctr[C_JOBPRIMS];
  for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < P; t += gridDim.x * kBlock) {
    const uint64_t k = rk[t];
    const uint32_t j = (uint32_t)(k >> 32);
    const uint32_t pos = c.ranges[c.jobs[j]].x + (t - c.jobStart[j]);
    keys[pos] = (uint32_t)k;
    c.ids[pos] = rv[t];
  }
}

// Thread phase (threadPhaseCreateNodes, :773-880): one thread finishes a
// range of <= 32 prims.  A range that cannot split is re-keyed in its own
// domain (or by local index) and sorted with the reference's exchange sort
// (bubbleSort, :758-771), whose swap sequence fixes the order of equal
// keys; if it still cannot split it becomes a leaf.  Local nodes (local 0 =
// the range's node) go to loc[2*begin + j]; boxes are computed bottom-up.
__global__ __launch_bounds__(kBlock) void k_thread_phase(RCtx c, uint32_t *keys) {
  const uint32_t nSmall = c.ctr[C_SMALL];
  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < nSmall; s += gridDim.x * kBlock) {
    const uint2 r = c.ranges[c.small[s]];
    Node *loc = c.loc + 2 * (size_t)r.x;
    uint32_t stB[kThreadPhase + 1], stE[kThreadPhase + 1], stM[kThreadPhase + 1];
    int sp = 0;
    stB[0] = r.x; stE[0] = r.y; stM[0] = 0; sp = 1;
    uint32_t nLocal = 1;
    while (sp > 0) {
      --sp;
      const uint32_t b = stB[sp], e = stE[sp], me = stM[sp];
      const uint32_t n = e - b;
      bool leaf = n <= c.T;
      uint32_t split = 0;
      if (!leaf) {
        if (keys[b] != keys[e - 1]) {
          split = find_split<uint32_t>(keys, b, e);
        } else {
          float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
          for (uint32_t i = b; i < e; i++) {
            float ce[3];
            center3(load_box(c.boxes, c.ids[i]), ce);
            for (int d = 0; d < 3; d++) { lo[d] = fminf(lo[d], ce[d]); hi[d] = fmaxf(hi[d], ce[d]); }
          }
          const bool point = lo[0] == hi[0] && lo[1] == hi[1] && lo[2] == hi[2];
          float q[4];
          quant_init(lo, hi, q);
          for (uint32_t i = b; i < e; i++) {
            if (point) {
              keys[i] = i - b;
            } else {
              float ce[3];
              center3(load_box(c.boxes, c.ids[i]), ce);
              keys[i] = morton<uint32_t>(ce, q, 10);
            }
          }
          for (uint32_t i = b; i < e; i++)
            for (uint32_t j = i + 1; j < e; j++) {
              const uint32_t ki = keys[i], kj = keys[j];
              if (kj < ki) {
                keys[i] = kj; keys[j] = ki;
                const uint32_t t = c.ids[i];
                c.ids[i] = c.ids[j];
                c.ids[j] = t;
              }
            }
          if (keys[b] != keys[e - 1]) split = find_split<uint32_t>(keys, b, e);
          else leaf = true;
        }
      }
      if (leaf) {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = b; i < e; i++) {
          const cubql_box3f bx = load_box(c.boxes, c.ids[i]);
          for (int d = 0; d < 3; d++) { lo[d] = fminf(lo[d], bx.lower[d]); hi[d] = fmaxf(hi[d], bx.upper[d]); }
        }
        Node nd;
        for (int d = 0; d < 3; d++) { nd.lo[d] = lo[d]; nd.hi[d] = hi[d]; }
        nd.oc = pack_oc(b, n);
        loc[me] = nd;
        continue;
      }
      const uint32_t child = nLocal;
      nLocal += 2;
      loc[me].oc = pack_oc(child, 0);
      stB[sp] = split; stE[sp] = e; stM[sp] = child + 1;
      stB[sp + 1] = b; stE[sp + 1] = split; stM[sp + 1] = child;
      sp += 2;
    }
    // children have larger local indices than their parent
    for (int j = (int)nLocal - 1; j >= 0; j--) {
      Node nd = loc[j];
      if (oc_count(nd.oc) == 0) {
        const Node l = loc[oc_offset(nd.oc)], rr = loc[oc_offset(nd.oc) + 1];
        for (int d = 0; d < 3; d++) { nd.lo[d] = fminf(l.lo[d], rr.lo[d]); nd.hi[d] = fmaxf(l.hi[d], rr.hi[d]); }
        loc[j] = nd;
      }
    }
    c.smallCount[s] = nLocal;
  }
}

// global ids of the thread-phase ranges' local nodes (the final node count
// goes to C_NODESOUT: every workgroup reads C_NODECOUNT as its base)
__global__ __launch_bounds__(1024) void k_thread_scan(RCtx c) {
  __shared__ uint32_t ss[1024 / kWave + 2];
  count_scan_tile(c.smallCount, c.smallBase, c.ctr[C_SMALL], c.ctr[C_NODECOUNT], c.scanStatus, 1u,
                  &c.ctr[C_NODESOUT], ss);
}

__global__ __launch_bounds__(kWave) void k_thread_emit(RCtx c, Node *out) {
  const uint32_t nSmall = c.ctr[C_SMALL];
  for (uint32_t s = blockIdx.x; s < nSmall; s += gridDim.x) {
    const uint32_t node = c.small[s];
    copy_local_nodes(c.loc + 2 * (size_t)c.ranges[node].x, c.smallCount[s], node, c.smallBase[s], out);
  }
}

int rebin_build(cubql_bvh3f *bvh, const cubql_box3f *boxes, uint32_t n, uint32_t T, hipStream_t s,
                const cubql_mem_resource *mem) {
  ScratchArena scr(mem, s);
  RCtx c{};
  c.boxes = boxes;
  c.n = n;
  c.T = T;
  c.bitsPerDim = 10;
  const uint32_t numBlocks = div_up(n, kTile);
  const uint32_t maxJobs = n / (kThreadPhase + 1) + 2;
  uint32_t *keysA = nullptr, *keysB = nullptr, *idsB = nullptr;
  uint64_t *rkA = nullptr, *rkB = nullptr;
  uint32_t *rvA = nullptr, *rvB = nullptr;
  CUBQL_CHECK(scr.reserve(&c.blockCount, numBlocks + 1));
  CUBQL_CHECK(scr.reserve(&c.blockBounds, 8 * (size_t)numBlocks + 8));
  CUBQL_CHECK(scr.reserve(&c.ctr, C_LSTART0 + kMaxLevels + 8));
  CUBQL_CHECK(scr.reserve(&c.quant, 8));
  CUBQL_CHECK(scr.reserve(&keysA, n));
  CUBQL_CHECK(scr.reserve(&keysB, n));
  CUBQL_CHECK(scr.reserve(&idsB, n));
  CUBQL_CHECK(scr.reserve(&c.nodes, 2 * (size_t)n + 2));
  CUBQL_CHECK(scr.reserve(&c.ranges, 2 * (size_t)n + 2));
  // a job node is listed again in the level after its rebin round
  CUBQL_CHECK(scr.reserve(&c.levels, 2 * (size_t)n + 2 + 64 * (size_t)maxJobs));
  CUBQL_CHECK(scr.reserve(&c.jobs, maxJobs));
  CUBQL_CHECK(scr.reserve(&c.jobStart, maxJobs + 1));
  CUBQL_CHECK(scr.reserve(&c.jobDom, 6 * (size_t)maxJobs));
  CUBQL_CHECK(scr.reserve(&c.jobQuant, 4 * (size_t)maxJobs));
  CUBQL_CHECK(scr.reserve(&c.jobLocal, maxJobs));
  CUBQL_CHECK(scr.reserve(&c.small, (size_t)n + 1));
  CUBQL_CHECK(scr.reserve(&c.smallCount, (size_t)n + 1));
  CUBQL_CHECK(scr.reserve(&c.smallBase, (size_t)n + 1));
  CUBQL_CHECK(scr.reserve(&c.loc, 2 * (size_t)n + 64));
  const uint32_t numScanTiles = div_up(n, kCountScanTile) + 1;
  CUBQL_CHECK(scr.reserve(&c.scanStatus, numScanTiles));
  CUBQL_CHECK(scr.reserve(&rkA, n));
  CUBQL_CHECK(scr.reserve(&rkB, n));
  CUBQL_CHECK(scr.reserve(&rvA, n));
  CUBQL_CHECK(scr.reserve(&rvB, n));
  CUBQL_CHECK(scr.commit());
  uint32_t *primIDs = nullptr;
  CUBQL_CHECK(mem_alloc(mem, (void **)&primIDs, (size_t)n * sizeof(uint32_t), s));
  c.ids = primIDs;
  auto fail = [&](int st) {
    mem_free(mem, primIDs, s);
    return st;
  };
  k_rinit_levels<<<1, 256, 0, s>>>(c.ctr);
  if (int fs = fill_u32((uint32_t *)c.scanStatus, 0u, 2 * (size_t)numScanTiles, s)) return fail(fs);
  int ks = prof_begin(s);
  k_rcount<<<numBlocks, kBlock, 0, s>>>(c);
  k_rscan<<<1, 1024, 0, s>>>(c, numBlocks);
  k_rkeys<uint32_t><<<numBlocks, kBlock, 0, s>>>(c, keysA);
  prof_end(ks, s, "rebin_keys");
  prof_set(ks, 2.0 * 24.0 * n + 8.0 * n, n);
  if (hipGetLastError() != hipSuccess) return fail((set_error("rebin key launch failed"), CUBQL_ERR_HIP));
  uint32_t *keys = nullptr, *vRes = nullptr;
  int st = radix_sort_pairs_dev_u32(keysA, primIDs, keysB, idsB, &c.ctr[C_NUMVALID], n, 30, s, mem, &keys, &vRes);
  if (st) return fail(st);
  if (vRes != primIDs &&
      hipMemcpyAsync(primIDs, vRes, (size_t)n * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return fail((set_error("id copy failed"), CUBQL_ERR_HIP));
  int dev = 0, numCU = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&numCU, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    numCU = 256;
  const uint32_t grid = std::max<uint32_t>(1u, std::min<uint32_t>(div_up(n, kBlock), (uint32_t)numCU * 8));
  int jobBits = 1;
  while ((1ull << jobBits) < (uint64_t)maxJobs) jobBits++;
  std::vector<uint32_t> hc(C_LSTART0 + kMaxLevels + 8);
  int level = 0, rounds = 0;
  int hs = prof_begin(s);
  while (true) {
    const int first = level;
    for (int i = 0; i < kLevelsPerSync && level < kMaxLevels; i++, level++)
      k_rlevel_rebin<<<grid, kBlock, 0, s>>>(c, keys, level);
    if (hipGetLastError() != hipSuccess) return fail((set_error("rebin level launch failed"), CUBQL_ERR_HIP));
    if (hipMemcpyAsync(hc.data(), c.ctr, sizeof(uint32_t) * hc.size(), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return fail((set_error("rebin level sync failed"), CUBQL_ERR_HIP));
    int L = -1;
    for (int l = first; l <= level; l++)
      if (hc[C_LEVEL0 + l] == 0) { L = l; break; }
    if (L < 0) {
      if (level >= kMaxLevels) return fail((set_error("rebin hierarchy deeper than %d levels", kMaxLevels), CUBQL_ERR_INTERNAL));
      continue;
    }
    if (hc[C_JOBS] == 0) { level = L; break; }
    if (L >= kMaxLevels - 1) return fail((set_error("rebin hierarchy deeper than %d levels", kMaxLevels), CUBQL_ERR_INTERNAL));
    // rebin round: the jobs become level L's open nodes with rebuilt keys
    rounds++;
    const uint32_t J = hc[C_JOBS];
    k_rebin_setup<<<1, 1024, 0, s>>>(c, L);
    k_rebin_grow<<<grid, kBlock, 0, s>>>(c);
    k_rebin_finalize<<<std::max<uint32_t>(1u, std::min<uint32_t>(div_up(J, kBlock), 1024u)), kBlock, 0, s>>>(c);
    k_rebin_keys<<<grid, kBlock, 0, s>>>(c, rkA, rvA);
    uint64_t *rkRes = nullptr;
    uint32_t *rvRes = nullptr;
    st = radix_sort_pairs_dev_u64(rkA, rvA, rkB, rvB, &c.ctr[C_JOBPRIMS], n, 32 + jobBits, s, mem, &rkRes, &rvRes);
    if (st) return fail(st);
    k_rebin_writeback<<<grid, kBlock, 0, s>>>(c, keys, rkRes, rvRes);
    if (hipGetLastError() != hipSuccess) return fail((set_error("rebin round launch failed"), CUBQL_ERR_HIP));
    level = L;
  }
  prof_end(hs, s, "rebin_hierarchy");
  const uint32_t numSmall = hc[C_SMALL];
  if (numSmall) {
    k_thread_phase<<<std::max<uint32_t>(1u, div_up(numSmall, kBlock)), kBlock, 0, s>>>(c, keys);
    k_thread_scan<<<div_up(numSmall, kCountScanTile), 1024, 0, s>>>(c);
  }
  if (hipGetLastError() != hipSuccess) return fail((set_error("thread phase launch failed"), CUBQL_ERR_HIP));
  if (hipMemcpyAsync(hc.data(), c.ctr, sizeof(uint32_t) * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail((set_error("rebin sync failed"), CUBQL_ERR_HIP));
  const uint32_t numValid = hc[C_NUMVALID];
  const uint32_t numNodes = numValid ? (numSmall ? hc[C_NODESOUT] : hc[C_NODECOUNT]) : 0;
  prof_set(hs, 8.0 * numNodes, numNodes);
  Node *nodes = nullptr;
  st = mem_alloc(mem, (void **)&nodes, std::max<size_t>(numNodes, 2) * sizeof(Node), s);
  if (st) return fail(st);
  if (numNodes) {
    if (hipMemcpyAsync(nodes, c.nodes, (size_t)numNodes * sizeof(Node), hipMemcpyDeviceToDevice, s) != hipSuccess) {
      mem_free(mem, nodes, s);
      return fail((set_error("node copy failed"), CUBQL_ERR_HIP));
    }
    if (numSmall) k_thread_emit<<<numSmall, kWave, 0, s>>>(c, nodes);
  }
  bvh->nodes = (cubql_node3f *)nodes;
  bvh->numNodes = numNodes;
  bvh->primIDs = primIDs;
  bvh->numPrims = numValid;
  int rs = prof_begin(s);
  refit_launch(c, nodes, primIDs, level, hc, numCU, s);
  st = hipGetLastError() == hipSuccess ? CUBQL_OK : (set_error("refit launch failed"), CUBQL_ERR_HIP);
  prof_end(rs, s, "refit");
  prof_set(rs, 24.0 * numValid + 64.0 * numNodes, numNodes);
  if (prof_on()) {
    if (hipStreamSynchronize(s) == hipSuccess) prof_flush();
  }
  (void)rounds;
  return st;
}

}  // namespace rb
}  // namespace cubql

using namespace cubql;

static int radix_entry(const char *fn, cubql_bvh3f *bvh, const cubql_box3f *boxes, uint32_t n,
                       const cubql_build_config *cfg, cubql_stream_t stream,
                       const cubql_mem_resource *mem, int bits) {
  clear_error();
  if (!bvh || !cfg || (n && !boxes)) {
    set_error("%s: null argument", fn);
    return CUBQL_ERR_INVALID_ARG;
  }
  if (n > 0x7fffffffu) {
    set_error("%s: too many prims", fn);
    return CUBQL_ERR_INVALID_ARG;
  }
  if (n == 0) return CUBQL_OK;
  // radix.h:596-600: threshold = min(makeLeafThreshold, maxAllowedLeafSize), 0 -> 1
  int T = cfg->makeLeafThreshold > 0 ? std::min(cfg->makeLeafThreshold, cfg->maxAllowedLeafSize) : 1;
  T = std::max(1, std::min(T, 65535));
  hipStream_t s = (hipStream_t)stream;
  if (bits == 21) return rb::build<uint64_t>(bvh, boxes, n, (uint32_t)T, 21, s, mem);
  return rb::build<uint32_t>(bvh, boxes, n, (uint32_t)T, 10, s, mem);
}

extern "C" int cubql_radix_builder_f3(cubql_bvh3f *bvh, const cubql_box3f *d_boxes, uint32_t numPrims,
                                      const cubql_build_config *cfg, cubql_stream_t s,
                                      const cubql_mem_resource *mem) {
  return radix_entry("cubql_radix_builder_f3", bvh, d_boxes, numPrims, cfg, s, mem, 21);
}

extern "C" int cubql_rebin_radix_builder_f3(cubql_bvh3f *bvh, const cubql_box3f *d_boxes,
                                            uint32_t numPrims, const cubql_build_config *cfg,
                                            cubql_stream_t stream, const cubql_mem_resource *mem) {
  clear_error();
  if (!bvh || !cfg || (numPrims && !d_boxes)) {
    set_error("cubql_rebin_radix_builder_f3: null argument");
    return CUBQL_ERR_INVALID_ARG;
  }
  if (numPrims > 0x7fffffffu) {
    set_error("cubql_rebin_radix_builder_f3: too many prims");
    return CUBQL_ERR_INVALID_ARG;
  }
  if (numPrims == 0) return CUBQL_OK;
  // rebinMortonBuilder.h:1175-1178
  int T = cfg->makeLeafThreshold > 0 ? std::min(cfg->makeLeafThreshold, cfg->maxAllowedLeafSize) : 1;
  T = std::max(1, std::min(T, 65535));
  return rb::rebin_build(bvh, d_boxes, numPrims, (uint32_t)T, (hipStream_t)stream, mem);
}

extern "C" int cubql_radix30_builder_f3(cubql_bvh3f *bvh, const cubql_box3f *d_boxes, uint32_t numPrims,
                                        const cubql_build_config *cfg, cubql_stream_t s,
                                        const cubql_mem_resource *mem) {
  return radix_entry("cubql_radix30_builder_f3", bvh, d_boxes, numPrims, cfg, s, mem, 10);
}
