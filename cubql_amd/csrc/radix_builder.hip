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
  C_SPIN = 7,      // a look-back wait exceeded its bound (build_common.h)
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

// ------------------------------------------------ one-pass hierarchy
// The tree findSplit/createNodes build top down (radix.h:446-570) is the
// binary radix tree of the sorted keys with equal keys told apart by their
// position, cut at ranges of <= T prims (leaves) and with every range of
// equal keys split in the middle instead (radix.h:455-465).  Each internal
// node of the radix tree is found independently from the keys around its
// position (Karras, "Maximizing parallelism in the construction of BVHs,
// octrees, and k-d trees", 2012): its range [first, last] and split gamma
// (left = [first, gamma], right = [gamma + 1, last]).  A node whose range
// has more than T prims and two different keys is a split node of the
// reference's tree; a child range of equal keys with more than T prims
// roots a middle-split subtree of mid_inner(size) inner nodes.
//
// Numbering: node 0 is the root and node 1 unused (radix.h:585-586); the
// children of the s-th split node (in position order) are nodes 2 + 2s and
// 3 + 2s; the pairs of the middle-split subtrees follow all of those, each
// subtree's in level order.  Boxes are refitted bottom up in buckets: the
// inner nodes of middle-split subtrees first, deepest first, then the split
// nodes in decreasing order of the common key-prefix length of their range,
// which grows strictly from a node to its split children, so every child
// is final before its parent.

// Readers of the sorted keys for the radix-tree search: from the
// workgroup's LDS copy of positions [lo, hi) (a read outside sets `miss`
// and returns 0; the node is then redone from global memory), or from
// global memory.  Two separate code paths keep LDS reads ds_ and global
// reads global_ loads (a merged read is a slower flat load).
template <typename K>
struct LdsKeys {
  const K *w;
  int64_t lo, hi;
  bool miss = false;
  template <typename I>
  __device__ __forceinline__ K at(I j) {
    if (j < lo || j >= hi) {
      miss = true;
      return (K)0;
    }
    return w[j - lo];
  }
};
template <typename K>
struct GlobalKeys {
  const K *g;
  template <typename I>
  __device__ __forceinline__ K at(I j) { return g[j]; }
};

// Common prefix length of key ki (at position i) and the key at j,
// positions breaking ties of equal keys (-1 out of range).  I: the index
// type (32-bit when n < 2^30, so that i + 2 * range never overflows).
template <typename K, typename I, typename R>
__device__ __forceinline__ int key_delta(R &rd, I n, I i, K ki, I j) {
  if (j < 0 || j >= n) return -1;
  const K b = rd.at(j);
  constexpr int kBits = 8 * (int)sizeof(K);
  if (ki == b) return kBits + __clz((uint32_t)(i ^ j));
  return sizeof(K) == 8 ? __clzll((unsigned long long)(ki ^ b)) : __clz((uint32_t)(ki ^ b));
}

// Range [first, last] and split gamma of radix-tree node i (Karras 2012),
// dn = the prefix length of the range.
template <typename K, typename I, typename R>
__device__ __forceinline__ void karras_node(R &rd, I n, I i, uint32_t &first, uint32_t &last, uint32_t &gamma,
                                            int &dn) {
  const K ki = rd.at(i);
  const int d = key_delta<K, I>(rd, n, i, ki, i + 1) - key_delta<K, I>(rd, n, i, ki, i - 1) > 0 ? 1 : -1;
  const int dmin = key_delta<K, I>(rd, n, i, ki, i - d);
  I lmax = 2;
  while (key_delta<K, I>(rd, n, i, ki, i + lmax * d) > dmin) lmax *= 2;
  I l = 0;
  for (I t = lmax / 2; t >= 1; t /= 2)
    if (key_delta<K, I>(rd, n, i, ki, i + (l + t) * d) > dmin) l += t;
  const I j = i + l * d;
  dn = key_delta<K, I>(rd, n, i, ki, j);
  I sp = 0;
  for (int sh = 1;; sh++) {
    const I t = (l + ((I)1 << sh) - 1) >> sh;  // ceil(l / 2^sh)
    if (key_delta<K, I>(rd, n, i, ki, i + (sp + t) * d) > dn) sp += t;
    if (t == 1) break;
  }
  gamma = (uint32_t)(i + sp * d + (d < 0 ? -1 : 0));
  first = (uint32_t)(i < j ? i : j);
  last = (uint32_t)(i < j ? j : i);
}

template <typename K>
__device__ __forceinline__ int key_prefix(K a, K b) {
  return sizeof(K) == 8 ? __clzll((unsigned long long)(a ^ b)) : __clz((uint32_t)(a ^ b));
}

// Shape of the middle-split subtree of m prims with leaves <= T (each inner
// node of s prims has children of floor(s/2) and ceil(s/2)): the sizes at
// depth d are q = m >> d or q + 1 (m mod 2^d of them), so every node above
// depth D = min{d : q_d <= T} is inner, at depth D exactly the nodes of
// T + 1 prims are (when q_D == T), and deeper nodes are leaves.
struct MidShape {
  uint32_t D, q, r, T;  // depth D, q_D, m mod 2^D, the leaf threshold
  // inner nodes at depth t
  __host__ __device__ uint32_t at(uint32_t t) const { return t < D ? 1u << t : (t == D && q == T) ? r : 0u; }
};

__host__ __device__ inline MidShape mid_shape(uint32_t m, uint32_t T) {
  MidShape ms{0u, m, 0u, T};
  while (ms.q > T) ms.q = m >> ++ms.D;
  ms.r = ms.D ? m & ((1u << ms.D) - 1u) : 0u;
  return ms;
}

// inner nodes of the middle-split subtree of m prims
__host__ __device__ inline uint32_t mid_inner(uint32_t m, uint32_t T) {
  const MidShape ms = mid_shape(m, T);
  return (ms.D ? (1u << ms.D) - 1u : 0u) + (ms.q == T ? ms.r : 0u);
}

// Refit buckets: split nodes by the prefix length of their range (< key
// bits), inner nodes of middle-split subtrees by key bits + 1 + depth; the
// last bin counts the middle-split subtree roots.
constexpr int kDeltaBins = 128;
constexpr int kBinEqRoots = kDeltaBins - 1;
// radix-tree nodes per workgroup of the hierarchy kernels (kKThreads
// threads, kKPer nodes each, strided); bucket slots are assigned per such
// block
constexpr int kKThreads = 1024;
constexpr int kKPer = 2;
constexpr uint32_t kKBlock = kKThreads * kKPer;

struct KNode {
  uint32_t first, last, gamma, pad;
};

// An equal-key range of more than T prims: the root of a middle-split
// subtree at node `pos`, prims [begin, begin + size), inner-node pairs from
// node `pairBase` on.
struct EqRoot {
  uint32_t pos, begin, size, pairBase;
};

// box of sorted prims [b, b + m): their ids are contiguous, their boxes
// are gathered (eight in flight at a time; for input in random order each
// box costs a line read)
__device__ __forceinline__ void range_box(const cubql_box3f *boxes, const uint32_t *primIDs, uint32_t b,
                                          uint32_t m, Node &nd) {
  for (int d = 0; d < 3; d++) { nd.lo[d] = INFINITY; nd.hi[d] = -INFINITY; }
  for (uint32_t j0 = 0; j0 < m; j0 += 8) {
    uint32_t id[8];
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (j0 + u < m) id[u] = primIDs[b + j0 + u];
    cubql_box3f bx[8];
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (j0 + u < m) bx[u] = load_box(boxes, id[u]);
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (j0 + u < m)
        for (int d = 0; d < 3; d++) {
          nd.lo[d] = fminf(nd.lo[d], bx[u].lower[d]);
          nd.hi[d] = fmaxf(nd.hi[d], bx[u].upper[d]);
        }
  }
}

// The per-node outputs of the radix-tree pass: the node record, whether it
// is a split node (more than T prims, two different keys), the inner nodes
// of the middle-split subtrees under it, and its refit-bucket counts
// (add(bin, count): into LDS in the block kernel, into the block's row in
// global memory for the deferred nodes).
template <typename K, typename Add>
__device__ __forceinline__ void karras_emit_counts(const K *keys, uint32_t T, uint32_t i, uint32_t first,
                                                   uint32_t last, uint32_t g, int dn, KNode *kn,
                                                   uint32_t *splitFlag, uint32_t *eqPairs, Add &&add) {
  constexpr int kBits = 8 * (int)sizeof(K);
  kn[i] = KNode{first, last, g, 0u};
  const bool split = last - first + 1 > T && keys[first] != keys[last];
  splitFlag[i] = split ? 1u : 0u;
  uint32_t eq = 0;
  if (split) {
    // (split nodes: the keys of the range differ, so dn < key bits)
    add(dn, 1u);
    const uint32_t cb[2] = {first, g + 1}, ce[2] = {g, last};
    for (int side = 0; side < 2; side++) {
      const uint32_t m = ce[side] - cb[side] + 1;
      if (m > T && keys[cb[side]] == keys[ce[side]]) {
        const MidShape ms = mid_shape(m, T);
        for (uint32_t t = 0; t <= ms.D; t++)
          if (const uint32_t c = ms.at(t)) add(kBits + 1 + (int)t, c);
        eq += mid_inner(m, T);
        add(kBinEqRoots, 1u);
      }
    }
  }
  eqPairs[i] = eq;
}

// keys a workgroup of k_karras stages in LDS: its nodes and kKWin on each
// side.  A node whose search reads outside (one with a long range, about
// one in five hundred) is deferred to k_karras_far.
constexpr int kKWin = 1024;

template <typename K>
__global__ __launch_bounds__(kKThreads) void k_karras(const K *keys, const uint32_t *nPtr, uint32_t T, KNode *kn,
                                                   uint32_t *splitFlag, uint32_t *eqPairs, uint32_t *blockHist,
                                                   uint32_t numKBlocks, uint32_t *farList, uint32_t *farCount) {
  __shared__ uint32_t sh[kDeltaBins];
  __shared__ K sKeys[kKBlock + 2 * kKWin];
  const uint32_t n = *nPtr;
  const int64_t N = n;
  const int64_t b0 = (int64_t)blockIdx.x * kKBlock;
  const int64_t wlo = b0 - kKWin < 0 ? 0 : b0 - kKWin;
  const int64_t whi = b0 + kKBlock + kKWin > N ? N : b0 + kKBlock + kKWin;
  for (int b = threadIdx.x; b < kDeltaBins; b += kKThreads) sh[b] = 0;
  for (int64_t j = wlo + threadIdx.x; j < whi; j += kKThreads) sKeys[j - wlo] = keys[j];
  __syncthreads();
  for (int r = 0; r < kKPer; r++) {
    const int64_t i = b0 + r * kKThreads + threadIdx.x;
    if (i >= N - 1) break;
    uint32_t first, last, g;
    int dn;
    LdsKeys<K> lk{sKeys, wlo, whi};
    if (N < (1 << 30)) {
      karras_node<K, int32_t>(lk, (int32_t)N, (int32_t)i, first, last, g, dn);
    } else {
      karras_node<K, int64_t>(lk, N, i, first, last, g, dn);
    }
    // deferred nodes: one atomic per wave claims the wave's list slots
    const uint64_t far = __ballot(lk.miss);
    if (far) {
      const uint32_t leader = (uint32_t)__builtin_ctzll(far);
      uint32_t base = 0;
      if (lane_id() == leader) base = atomicAdd(farCount, (uint32_t)__popcll(far));
      base = (uint32_t)__shfl((int)base, (int)leader, kWave);
      if (lk.miss) farList[base + (uint32_t)__popcll(far & ((1ull << lane_id()) - 1ull))] = (uint32_t)i;
    }
    if (lk.miss) continue;
    karras_emit_counts<K>(keys, T, (uint32_t)i, first, last, g, dn, kn, splitFlag, eqPairs,
                          [&](int bin, uint32_t c) { atomicAdd(&sh[bin], c); });
  }
  __syncthreads();
  for (int b = threadIdx.x; b < kDeltaBins; b += kKThreads) blockHist[(size_t)b * numKBlocks + blockIdx.x] = sh[b];
}

// Largest t in [0, tmax] with delta(i, i + t d) > thr, for one wave (all
// lanes get it): delta falls as t grows (sorted keys), so the test is true
// on a prefix of t.  Lane L first tests t = 2^L, then the bracket is cut in
// 64 per round; a 2^24 range takes five rounds of one load per lane.
template <typename K>
__device__ __forceinline__ int64_t wave_last_true(const K *keys, int64_t n, int64_t i, K ki, int d, int thr,
                                                  int64_t tmax) {
  GlobalKeys<K> gk{keys};
  const uint32_t lane = lane_id();
  const int64_t t0 = lane < 62 ? ((int64_t)1 << lane) : (int64_t)1 << 62;
  const bool ok0 = lane < 62 && t0 <= tmax && key_delta<K, int64_t>(gk, n, i, ki, i + t0 * d) > thr;
  const uint32_t k = (uint32_t)__popcll(__ballot(ok0));
  if (k == 0) return 0;
  int64_t lo = (int64_t)1 << (k - 1);
  int64_t hi = k < 62 ? ((int64_t)1 << k) : tmax + 1;
  if (hi > tmax + 1) hi = tmax + 1;
  while (hi - lo > 1) {
    const int64_t step = (hi - lo + kWave - 1) / kWave;
    const int64_t t = lo + (int64_t)(lane + 1) * step;
    const bool ok = t < hi && key_delta<K, int64_t>(gk, n, i, ki, i + t * d) > thr;
    lo += (int64_t)__popcll(__ballot(ok)) * step;
    if (lo + step < hi) hi = lo + step;
  }
  return lo;
}

// The deferred nodes, one wave each, reading keys from global memory.
template <typename K>
__global__ __launch_bounds__(256) void k_karras_far(const K *keys, const uint32_t *nPtr, uint32_t T,
                                                    const uint32_t *farList, const uint32_t *farCount, KNode *kn,
                                                    uint32_t *splitFlag, uint32_t *eqPairs, uint32_t *blockHist,
                                                    uint32_t numKBlocks) {
  const int64_t N = *nPtr;
  const uint32_t cnt = *farCount;
  const uint32_t nw = gridDim.x * (256 / kWave);
  for (uint32_t q = (blockIdx.x * 256 + threadIdx.x) / kWave; q < cnt; q += nw) {
    const int64_t i = farList[q];
    GlobalKeys<K> gk{keys};
    const K ki = keys[i];
    const int d = key_delta<K, int64_t>(gk, N, i, ki, i + 1) - key_delta<K, int64_t>(gk, N, i, ki, i - 1) > 0 ? 1 : -1;
    const int dmin = key_delta<K, int64_t>(gk, N, i, ki, i - d);
    const int64_t l = wave_last_true<K>(keys, N, i, ki, d, dmin, d > 0 ? N - 1 - i : i);
    const int64_t j = i + l * d;
    const int dn = key_delta<K, int64_t>(gk, N, i, ki, j);
    const int64_t sp = wave_last_true<K>(keys, N, i, ki, d, dn, l);
    const uint32_t g = (uint32_t)(i + sp * d + (d < 0 ? -1 : 0));
    const uint32_t first = (uint32_t)(i < j ? i : j), last = (uint32_t)(i < j ? j : i);
    if (lane_id() == 0) {
      uint32_t *row = blockHist + (size_t)(i / kKBlock);
      karras_emit_counts<K>(keys, T, (uint32_t)i, first, last, g, dn, kn, splitFlag, eqPairs,
                            [&](int bin, uint32_t c) { atomicAdd(&row[(size_t)bin * numKBlocks], c); });
    }
  }
}

// One workgroup per bin: exclusive scan of the bin's row over the blocks
// (in place), the bin's total to binTotal[bin].
__global__ __launch_bounds__(1024) void k_karras_binscan(uint32_t *blockHist, uint32_t numKBlocks,
                                                         uint32_t *binTotal) {
  __shared__ uint32_t ss[1024 / kWave + 1];
  uint32_t *row = blockHist + (size_t)blockIdx.x * numKBlocks;
  uint32_t run = 0;
  for (uint32_t b0 = 0; b0 < numKBlocks; b0 += 1024) {
    const uint32_t k = b0 + threadIdx.x;
    const uint32_t v = k < numKBlocks ? row[k] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan<1024>(v, ss, &tot);
    if (k < numKBlocks) row[k] = run + ex;
    run += tot;
  }
  if (threadIdx.x == 0) binTotal[blockIdx.x] = run;
}

// Exclusive scans of two per-node counts over the n - 1 radix-tree nodes
// (n = *nPtr valid prims) in one pass: one block scan for both, then waves
// 0 and 1 run the two tile look-backs (status arrays stA, stB) at once.
// listA[oA[i]] = i where cA[i] is 1 (the split nodes in position order).
__global__ __launch_bounds__(1024) void k_karras_scan2(const uint32_t *cA, const uint32_t *cB, uint32_t *oA,
                                                       uint32_t *oB, uint32_t *listA, const uint32_t *nPtr,
                                                       uint64_t *stA, uint64_t *stB, uint32_t *totA,
                                                       uint32_t *totB, uint32_t *err) {
  __shared__ uint32_t smem[2 * (1024 / kWave + 1)];
  __shared__ uint32_t sPre[2];
  const uint32_t n = *nPtr;
  const uint32_t m = n ? n - 1 : 0u;
  const uint32_t i0 = blockIdx.x * kCountScanTile + threadIdx.x * kCountScanPer;
  uint32_t a[kCountScanPer], b[kCountScanPer];
  uint32_t v[2] = {0u, 0u};
  load_per(cA, i0, m, a);
  load_per(cB, i0, m, b);
#pragma unroll
  for (uint32_t j = 0; j < kCountScanPer; j++) {
    v[0] += a[j];
    v[1] += b[j];
  }
  uint32_t ex[2], tot[2];
  block_excl_scan_multi<1024, 2>(v, ex, tot, smem);
  const uint32_t w = threadIdx.x / kWave;
  if (w < 2) {
    const uint32_t t = w == 0 ? tot[0] : tot[1];
    const uint32_t p = lookback_exclusive_wave(w == 0 ? stA : stB, blockIdx.x, blockIdx.x, 1u, t, false, err);
    if (lane_id() == 0) {
      sPre[w] = p;
      if (blockIdx.x == gridDim.x - 1) *(w == 0 ? totA : totB) = p + t;
    }
  }
  __syncthreads();
  uint32_t oa = sPre[0] + ex[0], ob = sPre[1] + ex[1];
  uint32_t ra[kCountScanPer], rb[kCountScanPer];
#pragma unroll
  for (uint32_t j = 0; j < kCountScanPer; j++) {
    ra[j] = oa;
    rb[j] = ob;
    if (i0 + j < m && a[j]) listA[oa] = i0 + j;
    oa += a[j];
    ob += b[j];
  }
  store_per(oA, i0, m, ra);
  store_per(oB, i0, m, rb);
}

// info[3] = 1 when the root is not a split node: at most T valid prims, or
// all of their keys equal
template <typename K>
__global__ void k_karras_info(const K *keys, const uint32_t *nPtr, uint32_t T, uint32_t *info) {
  const uint32_t n = *nPtr;
  info[3] = (n <= T || keys[0] == keys[n - 1]) ? 1u : 0u;
}

// The root alone: one leaf of at most T prims (one workgroup) ...
__global__ __launch_bounds__(256) void k_karras_leaf_root(const cubql_box3f *boxes, const uint32_t *primIDs,
                                                          const uint32_t *nPtr, Node *nodes) {
  __shared__ float sh[6][256];
  const uint32_t n = *nPtr;
  float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (uint32_t j = threadIdx.x; j < n; j += 256) {
    const cubql_box3f bx = load_box(boxes, primIDs[j]);
    for (int d = 0; d < 3; d++) { v[d] = fminf(v[d], bx.lower[d]); v[3 + d] = fmaxf(v[3 + d], bx.upper[d]); }
  }
  for (int k = 0; k < 6; k++) sh[k][threadIdx.x] = v[k];
  __syncthreads();
  for (uint32_t w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int k = 0; k < 6; k++)
        sh[k][threadIdx.x] = k < 3 ? fminf(sh[k][threadIdx.x], sh[k][threadIdx.x + w])
                                   : fmaxf(sh[k][threadIdx.x], sh[k][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    Node nd;
    for (int d = 0; d < 3; d++) { nd.lo[d] = sh[d][0]; nd.hi[d] = sh[3 + d][0]; }
    nd.oc = pack_oc(0, n);
    nodes[0] = nd;
  }
}

// ... or one middle-split subtree over all prims
__global__ void k_karras_root_eq(const uint32_t *nPtr, EqRoot *roots) { roots[0] = EqRoot{0u, 0u, *nPtr, 2u}; }

// Children of every split node (one thread per split node, from the list
// of split nodes): leaves with their boxes, split nodes (their position),
// roots of middle-split subtrees (claimed with one atomic per wave).  Each
// split node's refit bucket is its prefix length: dnKey[s] = dn, dnVal[s]
// = s, sorted by key afterwards.
template <typename K>
__global__ __launch_bounds__(kBlock) void k_karras_emit(const K *keys, const cubql_box3f *boxes,
                                                        const uint32_t *primIDs, uint32_t T,
                                                        const KNode *kn, const uint32_t *splitList,
                                                        uint32_t numSplit, const uint32_t *slot,
                                                        const uint32_t *eqPairs, const uint32_t *eqBase,
                                                        uint32_t *dnKey, uint32_t *dnVal, uint32_t *pos,
                                                        Node *nodes, EqRoot *eqRoots, uint32_t *eqRootCursor) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  const bool active = s < numSplit;
  uint32_t eqCount = 0;
  EqRoot er[2];
  if (active) {
    const uint32_t i = splitList[s];
    const KNode k = kn[i];
    const uint32_t base = 2 + 2 * s;
    if (s == 0) {
      // the root (node 0, the first split node)
      Node nd;
      for (int d = 0; d < 3; d++) { nd.lo[d] = INFINITY; nd.hi[d] = -INFINITY; }
      nd.oc = pack_oc(2, 0);
      nodes[0] = nd;
      pos[0] = 0;
    }
    dnKey[s] = (uint32_t)key_prefix(keys[k.first], keys[k.last]);
    dnVal[s] = s;
    uint32_t eqPair = eqPairs[i] ? 2 + 2 * numSplit + 2 * eqBase[i] : 0u;
    const uint32_t cb[2] = {k.first, k.gamma + 1}, ce[2] = {k.gamma, k.last};
    for (int side = 0; side < 2; side++) {
      const uint32_t m = ce[side] - cb[side] + 1;
      if (m <= T) {
        Node nd;
        range_box(boxes, primIDs, cb[side], m, nd);
        nd.oc = pack_oc(cb[side], m);
        nodes[base + side] = nd;
      } else if (keys[cb[side]] == keys[ce[side]]) {
        // equal keys: a middle-split subtree, built by k_karras_equal,
        // which also writes its root node
        er[eqCount++] = EqRoot{base + side, cb[side], m, eqPair};
        eqPair += 2 * mid_inner(m, T);
      } else {
        // the child range is the radix-tree node at its split end
        const uint32_t cs = slot[side == 0 ? k.gamma : k.gamma + 1];
        Node nd;
        for (int d = 0; d < 3; d++) { nd.lo[d] = INFINITY; nd.hi[d] = -INFINITY; }
        nd.oc = pack_oc(2 + 2 * cs, 0);
        nodes[base + side] = nd;
        pos[cs] = base + side;
      }
    }
  }
  // equal-key subtree roots: one atomic per wave claims the wave's slots
  const uint32_t incl = wave_incl_scan(eqCount);
  const uint32_t waveTot = (uint32_t)__shfl((int)incl, kWave - 1, kWave);
  if (waveTot) {
    uint32_t e0 = 0;
    if (lane_id() == kWave - 1) e0 = atomicAdd(eqRootCursor, waveTot);
    e0 = (uint32_t)__shfl((int)e0, kWave - 1, kWave) + incl - eqCount;
    for (uint32_t q = 0; q < eqCount; q++) eqRoots[e0 + q] = er[q];
  }
}

// Inner nodes of the middle-split subtrees.  Work item w of a subtree is
// the level-order inner node w when w < 2^D - 1, else node k = w - (2^D - 1)
// of depth D, inner when it holds T + 1 prims; the depth-D inner nodes are
// numbered by rank, which is their start offset minus k q (every depth-D
// node before k holds q or q + 1 prims).  Each inner node writes itself,
// its leaf children with their boxes, and its position into its depth's
// bucket (slots claimed by one atomic per wave and depth).
__global__ __launch_bounds__(256) void k_karras_equal(const cubql_box3f *boxes, const uint32_t *primIDs, uint32_t T, const EqRoot *roots, uint32_t nr, int kBits,
                                                      const uint32_t *bucketStart, uint32_t *bucketCursor,
                                                      uint32_t *bucket, Node *nodes) {
  const uint32_t lane = lane_id();
  for (uint32_t r = blockIdx.x; r < nr; r += gridDim.x) {
    const EqRoot er = roots[r];
    const MidShape ms = mid_shape(er.size, T);
    const uint32_t full = (1u << ms.D) - 1u;  // inner nodes above depth D
    const uint32_t items = full + (ms.q == T ? full + 1u : 0u);
    const uint32_t stride = gridDim.y * 256;
    // every lane of a wave runs the same number of rounds (the wave-wide
    // slot claim below needs all of them)
    for (uint32_t w0 = blockIdx.y * 256 + (threadIdx.x & ~(kWave - 1u)); w0 < items; w0 += stride) {
      const uint32_t w = w0 + lane;
      bool inner = false;
      uint32_t d = 0, pos = 0, pair = 0, b = 0, sz = 0;
      if (w < items) {
        uint32_t k;
        if (w < full) {
          d = 31u - (uint32_t)__clz(w + 1);
          k = w + 1 - (1u << d);
        } else {
          d = ms.D;
          k = w - full;
        }
        // range of node (d, k) from the root's halvings along k's bits
        sz = er.size;
        for (int j = (int)d - 1; j >= 0; j--) {
          const uint32_t h = sz >> 1;
          if ((k >> j) & 1u) {
            b += h;
            sz -= h;
          } else {
            sz = h;
          }
        }
        inner = sz > T;  // else a leaf at depth D, written by its parent
        if (inner) {
          const uint32_t I = d < ms.D ? w : full + (b - k * ms.q);
          pos = d == 0 ? er.pos : er.pairBase + 2 * ((1u << (d - 1)) - 1u + (k >> 1)) + (k & 1u);
          pair = er.pairBase + 2 * I;
          Node nd;
          for (int dd = 0; dd < 3; dd++) { nd.lo[dd] = INFINITY; nd.hi[dd] = -INFINITY; }
          nd.oc = pack_oc(pair, 0);
          nodes[pos] = nd;
          const uint32_t h = sz >> 1;
          const uint32_t cs[2] = {h, sz - h}, cbeg[2] = {b, b + h};
          for (int side = 0; side < 2; side++)
            if (cs[side] <= T) {
              Node lf;
              range_box(boxes, primIDs, er.begin + cbeg[side], cs[side], lf);
              lf.oc = pack_oc(er.begin + cbeg[side], cs[side]);
              nodes[pair + side] = lf;
            }
        }
      }
      // bucket slots: one atomic per distinct depth in the wave
      uint64_t rem = __ballot(inner);
      while (rem) {
        const uint32_t leader = (uint32_t)__builtin_ctzll(rem);
        const uint32_t dl = (uint32_t)__shfl((int)d, (int)leader, kWave);
        const uint64_t grp = __ballot(inner && d == dl);
        uint32_t slot0 = 0;
        if (lane == leader) slot0 = atomicAdd(&bucketCursor[kBits + 1 + dl], (uint32_t)__popcll(grp));
        slot0 = (uint32_t)__shfl((int)slot0, (int)leader, kWave);
        if (inner && d == dl)
          bucket[bucketStart[kBits + 1 + dl] + slot0 + (uint32_t)__popcll(grp & ((1ull << lane) - 1ull))] = pos;
        rem &= ~grp;
      }
    }
  }
}

// Refit of one bucket's inner nodes from their children, which are final
// (leaves carry their boxes; inner children were refitted in an earlier
// bucket).  Entries are split-node slots (node pos[s], children 2 + 2s) or,
// for middle-split buckets, node positions.
__device__ __forceinline__ void karras_refit_one(uint32_t e, bool slots, const uint32_t *pos, Node *nodes) {
  uint32_t p, c0;
  if (slots) {
    p = pos[e];
    c0 = 2 + 2 * e;
  } else {
    p = e;
    c0 = (uint32_t)oc_offset(nodes[p].oc);
  }
  const Node l = nodes[c0], r = nodes[c0 + 1];
  Node *me = &nodes[p];
  for (int d = 0; d < 3; d++) {
    me->lo[d] = fminf(l.lo[d], r.lo[d]);
    me->hi[d] = fmaxf(l.hi[d], r.hi[d]);
  }
}

__global__ __launch_bounds__(kBlock) void k_karras_refit(const uint32_t *list, uint32_t begin, uint32_t count,
                                                         bool slots, const uint32_t *pos, Node *nodes) {
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t < count) karras_refit_one(list[begin + t], slots, pos, nodes);
}

// The narrow buckets in one workgroup, in order; its barriers order them
// (one CU: a bucket's node writes are seen by the later buckets' reads).
// Split-node buckets (slots) are ranges of the slots sorted by prefix
// length; middle-split buckets are ranges of node positions.
struct RefitRange {
  uint32_t begin, count, slots, pad;
};
constexpr int kRefitTopBuckets = kDeltaBins;
__global__ __launch_bounds__(1024) void k_karras_refit_top(const uint32_t *slotList, const uint32_t *posList,
                                                           const RefitRange *ranges, int numRanges,
                                                           const uint32_t *pos, Node *nodes) {
  for (int r = 0; r < numRanges; r++) {
    const RefitRange rr = ranges[r];
    const uint32_t *list = rr.slots ? slotList : posList;
    for (uint32_t t = threadIdx.x; t < rr.count; t += 1024)
      karras_refit_one(list[rr.begin + t], rr.slots != 0, pos, nodes);
    __syncthreads();
  }
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
  // k_rscan's root writes (the hierarchy below does not use these)
  CUBQL_CHECK(scr.reserve(&c.nodes, 2));
  CUBQL_CHECK(scr.reserve(&c.ranges, 1));
  CUBQL_CHECK(scr.reserve(&c.levels, 1));
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
                                  n, keyBits, s, mem, (uint64_t **)&kRes, &vRes, &c.ctr[C_SPIN]);
  else
    st = radix_sort_pairs_dev_u32((uint32_t *)keysA, primIDs, (uint32_t *)keysB, idsB, &c.ctr[C_NUMVALID],
                                  n, keyBits, s, mem, (uint32_t **)&kRes, &vRes, &c.ctr[C_SPIN]);
  if (st) return fail(st);
  if (vRes != primIDs) {
    if (hipMemcpyAsync(primIDs, vRes, (size_t)n * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
      return fail((set_error("id copy failed"), CUBQL_ERR_HIP));
  }
  // hierarchy in one pass over the radix-tree nodes, one host check for the
  // node count, then the refit by prefix length
  int hs = prof_begin(s);
  constexpr int kBits = 8 * (int)sizeof(K);
  const uint32_t scanTiles = div_up(std::max(n, 1u), kCountScanTile);
  const uint32_t numKBlocks = div_up(std::max(n, 1u), kKBlock);
  enum { M_SPLIT = 0, M_EQPAIRS = 1, M_FAR = 2, M_ROOTEQ = 3, M_ERR = 4, M_EQCURSOR = 5, M_WORDS = 8 };
  KNode *kn = nullptr;
  uint32_t *splitFlag = nullptr, *slotArr = nullptr, *eqPairs = nullptr, *eqBase = nullptr, *misc = nullptr;
  uint32_t *hist = nullptr, *bStart = nullptr, *bCursor = nullptr, *bucket = nullptr, *pos = nullptr;
  uint32_t *splitList = nullptr, *dnKey = nullptr, *dnVal = nullptr, *dnKeyB = nullptr, *dnValB = nullptr;
  uint32_t *blockHist = nullptr;
  uint64_t *kStatus = nullptr;
  EqRoot *eqRoots = nullptr;
  RefitRange *topRanges = nullptr;
  ScratchArena hscr(mem, s);
  const size_t n1 = std::max(n, 1u);
  CUBQL_CHECK(hscr.reserve(&kn, n1));
  CUBQL_CHECK(hscr.reserve(&splitFlag, n1));
  CUBQL_CHECK(hscr.reserve(&slotArr, n1));
  CUBQL_CHECK(hscr.reserve(&eqPairs, n1));
  CUBQL_CHECK(hscr.reserve(&eqBase, n1));
  CUBQL_CHECK(hscr.reserve(&pos, n1));
  CUBQL_CHECK(hscr.reserve(&bucket, n1));  // middle-split inner nodes (< n)
  CUBQL_CHECK(hscr.reserve(&splitList, n1));
  CUBQL_CHECK(hscr.reserve(&dnKey, n1));
  CUBQL_CHECK(hscr.reserve(&dnVal, n1));
  CUBQL_CHECK(hscr.reserve(&dnKeyB, n1));
  CUBQL_CHECK(hscr.reserve(&dnValB, n1));
  CUBQL_CHECK(hscr.reserve(&eqRoots, n1));
  CUBQL_CHECK(hscr.reserve(&blockHist, (size_t)kDeltaBins * numKBlocks));
  CUBQL_CHECK(hscr.reserve(&kStatus, 2 * (size_t)scanTiles));
  CUBQL_CHECK(hscr.reserve(&misc, M_WORDS + 3 * kDeltaBins));
  CUBQL_CHECK(hscr.reserve(&topRanges, kRefitTopBuckets));
  uint32_t *farList = nullptr;
  CUBQL_CHECK(hscr.reserve(&farList, n1));
  if (hscr.commit() != CUBQL_OK) return fail(CUBQL_ERR_ALLOC);
  hist = misc + M_WORDS;
  bStart = hist + kDeltaBins;
  bCursor = bStart + kDeltaBins;
  if (fill_u32(misc, 0u, M_WORDS + 3 * kDeltaBins, s) != CUBQL_OK ||
      fill_u32((uint32_t *)kStatus, 0u, 4 * (size_t)scanTiles, s) != CUBQL_OK)
    return fail(CUBQL_ERR_HIP);
  // the valid count is on the device only: the grids are sized by n and
  // the kernels stop at the count
  const uint32_t *nv = &c.ctr[C_NUMVALID];
  k_karras_info<K><<<1, 1, 0, s>>>(kRes, nv, T, misc);
  k_karras<K><<<numKBlocks, kKThreads, 0, s>>>(kRes, nv, T, kn, splitFlag, eqPairs, blockHist, numKBlocks, farList,
                                               misc + M_FAR);
  k_karras_far<K><<<1024, 256, 0, s>>>(kRes, nv, T, farList, misc + M_FAR, kn, splitFlag, eqPairs, blockHist,
                                       numKBlocks);
  k_karras_binscan<<<kDeltaBins, 1024, 0, s>>>(blockHist, numKBlocks, hist);
  k_karras_scan2<<<scanTiles, 1024, 0, s>>>(splitFlag, eqPairs, slotArr, eqBase, splitList, nv, kStatus,
                                            kStatus + scanTiles,
                                            misc + M_SPLIT, misc + M_EQPAIRS, misc + M_ERR);
  if (hipGetLastError() != hipSuccess) return fail((set_error("hierarchy launch failed"), CUBQL_ERR_HIP));
  std::vector<uint32_t> hm(M_WORDS + kDeltaBins);
  uint32_t numValid = 0, spinWord = 0;
  if (hipMemcpyAsync(hm.data(), misc, sizeof(uint32_t) * hm.size(), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(&numValid, nv, sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(&spinWord, &c.ctr[C_SPIN], sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail((set_error("hierarchy sync failed"), CUBQL_ERR_HIP));
  prof_end(hs, s, "radix_hierarchy");
  prof_set(hs, 40.0 * numValid, numValid);
  if (hm[M_ERR] || spinWord)
    return fail((set_error("radix build: a look-back wait exceeded its bound"), CUBQL_ERR_INTERNAL));
  uint32_t *hh = hm.data() + M_WORDS;
  const bool rootEq = hm[M_ROOTEQ] != 0;
  const bool rootLeaf = numValid <= T;
  uint32_t numSplit = hm[M_SPLIT], eqInner = hm[M_EQPAIRS], numEqRoots = hh[kBinEqRoots];
  if (rootEq) {
    // no split node: the refit buckets are those of the one subtree
    numSplit = 0;
    numEqRoots = rootLeaf ? 0 : 1;
    eqInner = rootLeaf ? 0 : mid_inner(numValid, T);
    std::fill(hh, hh + kDeltaBins, 0u);
    if (!rootLeaf) {
      const MidShape ms = mid_shape(numValid, T);
      for (uint32_t t = 0; t <= ms.D; t++) hh[kBits + 1 + t] = ms.at(t);
    }
  }
  hh[kBinEqRoots] = 0;  // not a refit bucket
  const uint64_t nodes64 = numValid == 0 ? 0 : 2 + 2 * (uint64_t)numSplit + 2 * (uint64_t)eqInner;
  if (nodes64 > 0xffffffffull) return fail((set_error("radix build: node count overflows"), CUBQL_ERR_INTERNAL));
  const uint32_t numNodes = (uint32_t)nodes64;
  // bucket starts: split-node buckets are ranges of the slots sorted by
  // prefix length (ascending); middle-split buckets are ranges of `bucket`
  std::vector<uint32_t> hStart(kDeltaBins, 0);
  for (int d = 0, runS = 0, runE = 0; d < kDeltaBins; d++) {
    if (d < kBits) {
      hStart[d] = (uint32_t)runS;
      runS += (int)hh[d];
    } else {
      hStart[d] = (uint32_t)runE;
      runE += (int)hh[d];
    }
  }
  // refit order: middle-split buckets deepest first, then split-node
  // buckets by decreasing prefix length; the narrow tail (every bucket
  // after the last wide one) goes to one workgroup
  constexpr uint32_t kRefitTopMax = 4096;
  std::vector<RefitRange> order;
  for (int d = kDeltaBins - 1; d >= 0; d--)
    if (hh[d]) order.push_back(RefitRange{hStart[d], hh[d], d < kBits ? 1u : 0u, 0u});
  size_t topFrom = order.size();
  while (topFrom > 0 && order[topFrom - 1].count <= kRefitTopMax) topFrom--;
  Node *nodes = nullptr;
  int es = prof_begin(s);
  st = mem_alloc(mem, (void **)&nodes, std::max<size_t>(numNodes, 2) * sizeof(Node), s);
  if (st) return fail(st);
  auto failN = [&](int e) {
    mem_free(mem, nodes, s);
    return fail(e);
  };
  uint32_t *slotsByDn = dnVal;
  // node 1 stays unused (zero), as in the reference (radix.h:585-586)
  if (fill_u32((uint32_t *)(nodes + 1), 0u, sizeof(Node) / 4, s) != CUBQL_OK) return failN(CUBQL_ERR_HIP);
  if (numValid) {
    if (hipMemcpyAsync(bStart, hStart.data(), sizeof(uint32_t) * kDeltaBins, hipMemcpyHostToDevice, s) != hipSuccess ||
        (order.size() > topFrom &&
         hipMemcpyAsync(topRanges, order.data() + topFrom, sizeof(RefitRange) * (order.size() - topFrom),
                        hipMemcpyHostToDevice, s) != hipSuccess))
      return failN((set_error("bucket upload failed"), CUBQL_ERR_HIP));
    if (rootLeaf) {
      k_karras_leaf_root<<<1, 256, 0, s>>>(boxes, primIDs, nv, nodes);
    } else if (rootEq) {
      k_karras_root_eq<<<1, 1, 0, s>>>(nv, eqRoots);
    } else {
      k_karras_emit<K><<<div_up(numSplit, kBlock), kBlock, 0, s>>>(kRes, boxes, primIDs, T, kn, splitList, numSplit,
                                                                   slotArr, eqPairs, eqBase, dnKey, dnVal, pos,
                                                                   nodes, eqRoots, misc + M_EQCURSOR);
      // split nodes into prefix-length order (one 8-bit pass; lengths < 64).
      // No host check follows, so its look-back waits are the unbounded
      // kind (every tile waited on is held by a running workgroup).
      uint32_t *kOut = nullptr;
      st = radix_sort_pairs_dev_u32(dnKey, dnVal, dnKeyB, dnValB, misc + M_SPLIT, numSplit, 8, s, mem, &kOut,
                                    &slotsByDn, nullptr);
      if (st) return failN(st);
    }
    if (numEqRoots) {
      // a few workgroups per subtree when there are few subtrees
      const uint32_t gx = std::min(numEqRoots, 16384u), gy = std::min(64u, std::max(1u, 1024u / numEqRoots));
      k_karras_equal<<<dim3(gx, gy), 256, 0, s>>>(boxes, primIDs, T, eqRoots, numEqRoots, kBits, bStart,
                                                  bCursor, bucket, nodes);
    }
    if (hipGetLastError() != hipSuccess) return failN((set_error("hierarchy emit launch failed"), CUBQL_ERR_HIP));
  }
  prof_end(es, s, "radix_emit");
  prof_set(es, 32.0 * numNodes + 24.0 * numValid, numNodes);
  bvh->nodes = (cubql_node3f *)nodes;
  bvh->numNodes = numNodes;
  bvh->primIDs = primIDs;
  bvh->numPrims = numValid;
  int rs = prof_begin(s);
  if (numValid && !rootLeaf) {
    for (size_t r = 0; r < topFrom; r++)
      k_karras_refit<<<div_up(order[r].count, kBlock), kBlock, 0, s>>>(order[r].slots ? slotsByDn : bucket,
                                                                        order[r].begin, order[r].count,
                                                                        order[r].slots != 0, pos, nodes);
    if (order.size() > topFrom)
      k_karras_refit_top<<<1, 1024, 0, s>>>(slotsByDn, bucket, topRanges, (int)(order.size() - topFrom), pos,
                                            nodes);
  }
  st = hipGetLastError() == hipSuccess ? CUBQL_OK : (set_error("refit launch failed"), CUBQL_ERR_HIP);
  prof_end(rs, s, "refit");
  prof_set(rs, 96.0 * (numNodes / 2), numNodes / 2);
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

// Centroid domain of every job, over all rebin prims at once.  Each wave
// walks a contiguous run of 64-prim groups and keeps the bounds of its
// current job in registers while its groups stay in that job (jobs are
// contiguous and usually large), adding them to the job's domain (one
// atomic per bound) when the job changes and at the end of the run; a group
// that straddles jobs adds its prims lane by lane.  So a job takes about
// one atomic per wave and bound, not one per group: the domains of large
// jobs are the same few words, and same-address atomics serialise.
// jobDom holds encoded bounds.
__global__ __launch_bounds__(kBlock) void k_rebin_grow(RCtx c) {
  const uint32_t J = c.ctr[C_CURJOBS];
  const uint32_t P = c.ctr[C_JOBPRIMS];
  const uint32_t lane = lane_id();
  const uint32_t nWaves = gridDim.x * (kBlock / kWave);
  const uint32_t w = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  const uint32_t groups = div_up(P, (uint32_t)kWave);
  const uint32_t per = div_up(groups, nWaves);
  const uint32_t g0 = min(groups, w * per), g1 = min(groups, g0 + per);
  float acc[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  uint32_t accJob = 0xffffffffu;  // (wave-uniform)
  auto flush = [&]() {
    if (accJob == 0xffffffffu) return;
    float v[6];
    for (int d = 0; d < 3; d++) {
      v[d] = wave_min_dpp(acc[d]);
      v[d + 3] = wave_max_dpp(acc[d + 3]);
    }
    if (lane < 6) {
      const float x = lane == 0 ? v[0] : lane == 1 ? v[1] : lane == 2 ? v[2] : lane == 3 ? v[3] : lane == 4 ? v[4] : v[5];
      if (lane < 3) atomicMin(&c.jobDom[6 * (size_t)accJob + lane], enc_f(x));
      else atomicMax(&c.jobDom[6 * (size_t)accJob + lane], enc_f(x));
    }
  };
  for (uint32_t g = g0; g < g1; g++) {
    const uint32_t t = g * kWave + lane;
    const bool active = t < P;
    uint32_t j = 0;
    float ce[3] = {0.f, 0.f, 0.f};
    if (active) {
      j = find_job(c.jobStart, J, t);
      center3(load_box(c.boxes, c.ids[c.ranges[c.jobs[j]].x + (t - c.jobStart[j])]), ce);
    }
    const uint32_t j0 = __builtin_amdgcn_readfirstlane(j);  // lane 0 is active: g * 64 < P
    if (__ballot(active && j != j0) == 0) {
      if (j0 != accJob) {
        flush();
        for (int d = 0; d < 3; d++) { acc[d] = INFINITY; acc[d + 3] = -INFINITY; }
        accJob = j0;
      }
      for (int d = 0; d < 3; d++) {
        acc[d] = fminf(acc[d], active ? ce[d] : INFINITY);
        acc[d + 3] = fmaxf(acc[d + 3], active ? ce[d] : -INFINITY);
      }
    } else if (active) {
      int32_t *dom = c.jobDom + 6 * (size_t)j;
      for (int d = 0; d < 3; d++) {
        atomicMin(&dom[d], enc_f(ce[d]));
        atomicMax(&dom[d + 3], enc_f(ce[d]));
      }
    }
  }
  flush();
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
                  &c.ctr[C_NODESOUT], ss, 1u, &c.ctr[C_SPIN]);
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
  int st = radix_sort_pairs_dev_u32(keysA, primIDs, keysB, idsB, &c.ctr[C_NUMVALID], n, 30, s, mem, &keys, &vRes,
                                    &c.ctr[C_SPIN]);
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
    if (hc[C_SPIN]) return fail((set_error("rebin build: a look-back wait exceeded its bound"), CUBQL_ERR_INTERNAL));
    if (hc[C_JOBS] == 0) { level = L; break; }
    if (L >= kMaxLevels - 1) return fail((set_error("rebin hierarchy deeper than %d levels", kMaxLevels), CUBQL_ERR_INTERNAL));
    // rebin round: the jobs become level L's open nodes with rebuilt keys
    rounds++;
    const uint32_t J = hc[C_JOBS];
    // job ids are below J: the sort needs no more key bits than that
    int roundJobBits = 1;
    while ((1ull << roundJobBits) < (uint64_t)J) roundJobBits++;
    roundJobBits = std::min(roundJobBits, jobBits);
    k_rebin_setup<<<1, 1024, 0, s>>>(c, L);
    k_rebin_grow<<<grid, kBlock, 0, s>>>(c);
    k_rebin_finalize<<<std::max<uint32_t>(1u, std::min<uint32_t>(div_up(J, kBlock), 1024u)), kBlock, 0, s>>>(c);
    k_rebin_keys<<<grid, kBlock, 0, s>>>(c, rkA, rvA);
    uint64_t *rkRes = nullptr;
    uint32_t *rvRes = nullptr;
    st = radix_sort_pairs_dev_u64(rkA, rvA, rkB, rvB, &c.ctr[C_JOBPRIMS], n, 32 + roundJobBits, s, mem, &rkRes, &rvRes,
                                  &c.ctr[C_SPIN]);
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
  if (hc[C_SPIN]) return fail((set_error("rebin build: a look-back wait exceeded its bound"), CUBQL_ERR_INTERNAL));
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
