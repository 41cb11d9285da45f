// sah_builder.hip -- binned SAH BVH builder for gfx950.
//
// Replaces cuBQL::cuda::sahBuilder (builder/cuda/sah_builder.h:396-537) and
// gpuBuilder's SAH dispatch (gpu_builder.h:31-67).  Split rules are the
// reference's: per open node and per dimension, prims are binned by centroid
// (bin = clamp(int((c-lo)/(hi-lo+1e-20f)*numBins)), sah_builder.h:181-187);
// a bin accumulates the prim boxes and a count; the first minimum of [leaf
// cost A*N, then dim 0..2 x bin 0..numBins-1 split cost A_L*N_L + A_R*N_R]
// decides (evaluateSAH, :62-106), and a prim goes right iff its bin >= the
// split bin (:315-321).  Additions: numBins is a parameter (<= 64), and
// BuildConfig::maxAllowedLeafSize is honoured (a leaf choice above the cap
// takes the best split, or the middle of the node's range when no plane
// separates the prims), which also keeps counts inside the 16-bit field.
//
// MI355X design (not the reference's prim-state + global sort scheme):
//   * prims stay physically partitioned as 32-byte items {box, primID}, each
//     open node owning a contiguous range (left side stable, right side
//     reversed), so the tree and primIDs are a deterministic function of the
//     input;
//   * large nodes: binning tiles of 8192 prims build their bins in LDS and
//     write one partial per tile (no global atomics); one workgroup per node
//     reduces the partials and one wave evaluates all 3 x numBins candidate
//     planes at once (prefix/suffix unions as wave scans); the partition is
//     a single-pass decoupled look-back over 2048-prim tiles;
//   * nodes of <= 256 prims are finished by one wave each in LDS, which also
//     computes their boxes, so no refit pass is needed;
//   * the level loop runs on the device; the host checks for completion every
//     few levels.
#include "block_ops.h"
#include "build_common.h"
#include "common.h"
#include "profile.h"
#include "wave_ops.h"

#include <algorithm>
#include <vector>

namespace cubql {
namespace sah {

constexpr int kBlock = 256;
constexpr int kItems = 8;
constexpr int kTile = kBlock * kItems;  // partition tile: 2048 prims
constexpr int kBinTile = 4 * kTile;     // binning tile: 8192 prims
constexpr int kSmall = 256;             // max prims of a wave-finished node
constexpr int kMaxBins = 64;            // one lane per bin
constexpr int kBinWords = 8;            // lo xyz, hi xyz (encoded), count, pad
constexpr int kLevelsPerSync = 8;
// interleaved tile-ticket counters and child-bound copies spread the
// same-address atomics of the partition (see sm_builder.hip)
constexpr uint32_t kTicketCounters = 8;
constexpr uint32_t kTicketStride = 32;
constexpr uint32_t kCbCopies = 8;

enum : int32_t { kModeLeaf = -1, kModeMiddle = -2 };

struct alignas(16) Item {
  float4 a;  // lower.xyz, primID bits
  float4 b;  // upper.xyz, 0
};

struct Counters {
  uint32_t numValid;
  int32_t rootCb[6];
  uint32_t nodeCount;
  uint32_t numNodesOut;  // node count after the small phase
  uint32_t numLarge[2];
  uint32_t numSmall;
  uint32_t ticket;
  uint32_t numTiles;
  uint32_t numBinTiles;
  uint32_t spinFault;  // a look-back wait exceeded its bound (build_common.h)
  uint32_t pad;
};

struct Seg {
  uint32_t begin, end, node, pad;
  int32_t cb[6];  // encoded centroid bounds
  uint32_t pad2[2];
};

struct SmallSeg {
  uint32_t begin, end, node, buf;
};

struct Ctx {
  const cubql_box3f *boxes;
  uint32_t n;    // input boxes
  uint32_t T;    // effective leaf threshold (<= cap)
  uint32_t cap;  // leaf size cap
  int nb;        // bins per dimension
  float binEps;  // bin denominator epsilon: 1e-20 (SAH) or 0 (ELH)
  int elh;       // box measure: surface area (0) or edge-length sum (1)
  Item *items[2];
  uint32_t *primIDs;
  Node *nodes;   // level-phase nodes (capacity 2*n+2)
  Counters *ctr;
  uint32_t *blockCount;
  float *blockBounds;
  Seg *segs[2];
  int32_t *childCb;    // [kCbCopies][2*maxLarge][6]
  uint32_t *tickets;   // kTicketCounters counters, kTicketStride words apart
  uint32_t *segLeft;   // [maxLarge]
  uint32_t *tileStart; // [maxLarge+1] partition tiles
  uint32_t *binStart;  // [maxLarge+1] binning tiles
  int32_t *segMode;    // split dim, kModeLeaf or kModeMiddle
  int32_t *segBin;
  float *segLo, *segHi;  // centroid bounds of the split dim
  uint64_t *tileStatus;
  int32_t *partial;    // [binTiles][3*nb*kBinWords]
  SmallSeg *small;
  Node *loc;
  uint32_t *smallCount, *smallBase;
  uint32_t *levelItems;
  uint32_t maxLarge;
  uint32_t maxChunks;      // 1024-segment chunks of the emit
  uint64_t *emitStatus;    // [6][maxChunks] look-back words of the emit's counts
};

// The partition and binning tilings of a level's open segments: tile
// offsets per segment and the totals, and the partition's tile tickets
// cleared.  Written by the emit of the level before (k_count_scan for the
// root).
__device__ __forceinline__ void level_plan_totals(const Ctx &c, uint32_t level, uint32_t nSeg, uint32_t tiles,
                                                  uint32_t btiles, uint32_t items) {
  c.tileStart[nSeg] = tiles;
  c.binStart[nSeg] = btiles;
  c.ctr->numTiles = tiles;
  c.ctr->numBinTiles = btiles;
  c.ctr->ticket = 0;
  for (uint32_t cc = 0; cc < kTicketCounters; cc++) c.tickets[kTicketStride * cc] = 0;
  c.levelItems[level] = items;
}

__device__ __forceinline__ Item make_item(const cubql_box3f &b, uint32_t id) {
  Item it;
  it.a = make_float4(b.lower[0], b.lower[1], b.lower[2], __uint_as_float(id));
  it.b = make_float4(b.upper[0], b.upper[1], b.upper[2], 0.f);
  return it;
}
__device__ __forceinline__ float item_center(const Item &it, int d) {
  return d == 0 ? (it.a.x + it.b.x) / 2.f : (d == 1 ? (it.a.y + it.b.y) / 2.f : (it.a.z + it.b.z) / 2.f);
}
__device__ __forceinline__ uint32_t item_id(const Item &it) { return __float_as_uint(it.a.w); }

// bin of a centre coordinate (sah_builder.h:181-187, 301-308).  Clamping in
// float equals max(0,min(nb-1,int(rel*nb))) for in-range values and gives
// the saturating conversion's answer for NaN and infinities.
__device__ __forceinline__ int sah_bin(float c, float lo, float hi, int nb, float eps) {
  if (!(hi > lo)) return 0;
  const float rel = (c - lo) / (hi - lo + eps);
  const float f = rel * (float)nb;
  if (f >= (float)(nb - 1)) return nb - 1;
  if (f > 0.f) return (int)f;
  return 0;
}

// surfaceArea(box3f) (math/box.h:199-205), or for the edge-length
// heuristic edgeLengths(box) (elh_builder.h:66-72): the extents summed from
// 0 in dimension order
__device__ __forceinline__ float box_measure(const float (&lo)[3], const float (&hi)[3], bool elh) {
  const float sx = hi[0] - lo[0], sy = hi[1] - lo[1], sz = hi[2] - lo[2];
  return elh ? (sx + sy) + sz : sx * sy + sx * sz + sy * sz;
}

__device__ __forceinline__ int32_t bin_init(int w) {
  const int k = w & (kBinWords - 1);
  return k < 3 ? enc_f(INFINITY) : (k < 6 ? enc_f(-INFINITY) : 0);
}

__device__ __forceinline__ void bin_add(int32_t *bins, int nb, float eps, const Item &it, const float (&clo)[3],
                                        const float (&chi)[3]) {
#pragma unroll
  for (int d = 0; d < 3; d++) {
    int32_t *w = bins + (d * nb + sah_bin(item_center(it, d), clo[d], chi[d], nb, eps)) * kBinWords;
    atomicMin(&w[0], enc_f(it.a.x));
    atomicMin(&w[1], enc_f(it.a.y));
    atomicMin(&w[2], enc_f(it.a.z));
    atomicMax(&w[3], enc_f(it.b.x));
    atomicMax(&w[4], enc_f(it.b.y));
    atomicMax(&w[5], enc_f(it.b.z));
    atomicAdd(&w[6], 1);
  }
}

// Split costs for the candidates held by this lane's (dim, bin): the left
// side is the prefix scan over bins shifted by one lane, the right side the
// prefix scan over the reversed bins read back from lane rev.
// Returns +inf for invalid candidates; *leafTerm = area * count of all bins.
template <bool CROSS32>
__device__ __forceinline__ float split_cost(const int32_t *bins, int nb, int d, uint32_t bl, bool inBin,
                                            int rev, bool elh, float *leafTerm) {
  const bool hasLeft = bl >= 1u && bl <= (uint32_t)nb;
  const uint32_t binL = hasLeft ? bl - 1u : 0u;
  const uint32_t binR = inBin ? (uint32_t)nb - 1u - bl : 0u;
  const int32_t *wl = bins + (d * nb + binL) * kBinWords;
  const int32_t *wr = bins + (d * nb + binR) * kBinWords;
  float llo[3], lhi[3], rlo[3], rhi[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    llo[k] = dec_f(dpp_scan_i<0, CROSS32>(hasLeft ? wl[k] : INT32_MAX));
    lhi[k] = dec_f(dpp_scan_i<1, CROSS32>(hasLeft ? wl[3 + k] : INT32_MIN));
    rlo[k] = dec_f(dpp_scan_i<0, CROSS32>(inBin ? wr[k] : INT32_MAX));
    rhi[k] = dec_f(dpp_scan_i<1, CROSS32>(inBin ? wr[3 + k] : INT32_MIN));
  }
  const uint32_t lCnt = (uint32_t)dpp_scan_i<2, CROSS32>(hasLeft ? wl[6] : 0);
  const uint32_t rCntRev = (uint32_t)dpp_scan_i<2, CROSS32>(inBin ? wr[6] : 0);
  const float rArea = __shfl(box_measure(rlo, rhi, elh), rev, kWave);
  const uint32_t rCnt = (uint32_t)__shfl((int)rCntRev, rev, kWave);
  *leafTerm = rArea * (float)rCnt;
  const float c = box_measure(llo, lhi, elh) * (float)lCnt + rArea * (float)rCnt;
  return (inBin && lCnt > 0 && rCnt > 0 && c == c) ? c : INFINITY;
}

// evaluateSAH (sah_builder.h:62-106) for one node, by one full wave: each
// lane holds one (dim, bin) split candidate.  Every prefix/suffix union is
// exact (min/max), so costs equal the serial evaluation's bit for bit; the
// first minimum in (dim, bin) order is found by wave min + ballot.  With at
// most 32 bins dims 0 and 1 share one pass (lanes 0-31, 32-63).
// n = prims of the node.  Returns the split dim, kModeLeaf or kModeMiddle.
__device__ int sah_decide_wave(const int32_t *bins, int nb, uint32_t n, uint32_t cap, bool elh, int &splitBin) {
  const uint32_t lane = lane_id();
  float cost[3];     // candidate sets in (dim, bin) order; lane order within each
  int passDim[3];    // first dim of each candidate set (-1: unused)
  int nSets;
  float leafCost;
  if (nb <= 32) {
    const uint32_t bl = lane & 31u, half = lane >> 5;
    float leaf0, leaf1;
    cost[0] = split_cost<false>(bins, nb, (int)half, bl, bl < (uint32_t)nb, (int)(32 * half + nb - 1 - bl) & 63,
                                elh, &leaf0);
    cost[1] = split_cost<false>(bins, nb, 2, bl, half == 0 && bl < (uint32_t)nb, (int)(nb - 1 - bl) & 63, elh,
                                &leaf1);
    passDim[0] = 0; passDim[1] = 2; passDim[2] = -1;
    nSets = 2;
    leafCost = __shfl(leaf0, 0, kWave);
  } else {
    const bool inBin = lane < (uint32_t)nb;
    float lt[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      cost[d] = split_cost<true>(bins, nb, d, lane, inBin, inBin ? nb - 1 - (int)lane : (int)lane, elh, &lt[d]);
      passDim[d] = d;
    }
    nSets = 3;
    leafCost = __shfl(lt[0], 0, kWave);
  }
  // rAreas[0] * rCount: the union of all bins is the node box
  auto pick = [&](float thr, int &bin) -> int {
    float best = thr;
    int dim = kModeLeaf;
    for (int p = 0; p < nSets; p++) {
      // (costs are +0 .. +inf: their bit patterns order like the values)
      const float m = __int_as_float(wave_min_i32(__float_as_int(cost[p])));
      if (m < best) {
        best = m;
        const uint32_t idx = (uint32_t)__builtin_ctzll(__ballot(cost[p] == m));
        if (nb <= 32) {
          dim = passDim[p] + (int)(idx >> 5);
          bin = (int)(idx & 31u);
        } else {
          dim = passDim[p];
          bin = (int)idx;
        }
      }
    }
    return dim;
  };
  // the leaf candidate comes first: a split must beat it strictly
  int dim = pick(leafCost < INFINITY ? leafCost : INFINITY, splitBin);
  if (dim == kModeLeaf && n > cap) {
    dim = pick(INFINITY, splitBin);
    if (dim == kModeLeaf) dim = kModeMiddle;
  }
  return dim;
}

// ---------------------------------------------------------------- init pass
__global__ __launch_bounds__(kBlock) void k_count(Ctx c) {
  __shared__ float sb[12 * (kBlock / kWave)];
  __shared__ uint32_t ss[kBlock / kWave + 1];
  const uint32_t base = blockIdx.x * kTile;
  uint32_t cnt = 0;
  float v[12];
  for (int j = 0; j < 3; j++) { v[j] = INFINITY; v[j + 3] = -INFINITY; v[j + 6] = INFINITY; v[j + 9] = -INFINITY; }
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    const uint32_t i = base + k * kBlock + threadIdx.x;
    if (i < c.n) {
      const cubql_box3f b = load_box(c.boxes, i);
      if (box_is_valid(b)) {
        cnt++;
        for (int d = 0; d < 3; d++) {
          const float ce = (b.lower[d] + b.upper[d]) / 2.f;
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

__global__ __launch_bounds__(1024) void k_count_scan(Ctx c, uint32_t numBlocks) {
  __shared__ uint32_t ss[1024 / kWave + 1];
  __shared__ float sr[6][1024 / kWave];
  uint32_t run = 0;
  float red[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (uint32_t b0 = 0; b0 < numBlocks; b0 += 1024) {
    const uint32_t i = b0 + threadIdx.x;
    const uint32_t v = i < numBlocks ? c.blockCount[i] : 0;
    if (i < numBlocks) {
      const float *bb = c.blockBounds + (size_t)i * 8;
      for (int j = 0; j < 3; j++) {
        red[j] = fminf(red[j], bb[j]);
        red[j + 3] = fmaxf(red[j + 3], bb[j + 3]);
      }
    }
    uint32_t tot;
    const uint32_t ex = block_excl_scan<1024>(v, ss, &tot);
    if (i < numBlocks) c.blockCount[i] = run + ex;
    run += tot;
  }
  for (int j = 0; j < 6; j++) {
    const float r = j >= 3 ? wave_max(red[j]) : wave_min(red[j]);
    if (lane_id() == 0) sr[j][threadIdx.x / kWave] = r;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    Counters *k = c.ctr;
    for (int j = 0; j < 6; j++) {
      float r = sr[j][0];
      for (int w = 1; w < 1024 / kWave; w++) r = j >= 3 ? fmaxf(r, sr[j][w]) : fminf(r, sr[j][w]);
      k->rootCb[j] = enc_f(r);
    }
    k->numValid = run;
    k->spinFault = 0;
    k->nodeCount = 2;
    k->numLarge[0] = k->numLarge[1] = 0;
    k->numSmall = 0;
    Node z;
    for (int d = 0; d < 3; d++) { z.lo[d] = 0.f; z.hi[d] = 0.f; }
    z.oc = 0;
    c.nodes[1] = z;  // node 1 is never used (bvh.h:66-70)
    if (run > (uint32_t)kSmall && run > c.T) {
      Seg s;
      s.begin = 0; s.end = run; s.node = 0; s.pad = 0; s.pad2[0] = s.pad2[1] = 0;
      for (int j = 0; j < 6; j++) s.cb[j] = k->rootCb[j];
      c.segs[0][0] = s;
      k->numLarge[0] = 1;
      c.tileStart[0] = 0;
      c.binStart[0] = 0;
      level_plan_totals(c, 0u, 1u, div_up(run, kTile), div_up(run, kBinTile), run);
    } else {
      if (run > 0) {
        c.small[0] = SmallSeg{0, run, 0, 0};
        k->numSmall = 1;
      }
      level_plan_totals(c, 0u, 0u, 0u, 0u, 0u);
    }
  }
}

// compacted items in ascending primID order; invalid prims at the tail
__global__ __launch_bounds__(kBlock) void k_write_items(Ctx c) {
  __shared__ uint32_t sr[kItems * (kBlock / kWave) + 1];
  const uint32_t base = blockIdx.x * kTile;
  bool valid[kItems];
  cubql_box3f bx[kItems];
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    const uint32_t i = base + k * kBlock + threadIdx.x;
    valid[k] = false;
    if (i < c.n) {
      bx[k] = load_box(c.boxes, i);
      valid[k] = box_is_valid(bx[k]);
    }
  }
  uint32_t rank[kItems];
  block_rank_striped<kBlock, kItems>(valid, rank, sr);
  const uint32_t off = c.blockCount[blockIdx.x];
  const uint32_t numValid = c.ctr->numValid;
  const uint32_t invBefore = base - off;
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    const uint32_t i = base + k * kBlock + threadIdx.x;
    if (i >= c.n) continue;
    const uint32_t local = k * kBlock + threadIdx.x;
    if (valid[k]) c.items[0][off + rank[k]] = make_item(bx[k], i);
    else c.primIDs[numValid + invBefore + (local - rank[k])] = i;
  }
}

// ------------------------------------------------------------- level pass
// The emit runs one 1024-thread workgroup per 1024 open segments (a chunk;
// chunk ch is workgroup ch, so every chunk waited on is held by a workgroup
// the dispatcher started earlier): K counts per segment, block scans, and
// the chunks' offsets by decoupled look-back, one wave per count at once.
// The status words carry the level as their tag, so they are cleared once
// per build.  The chunk's exclusive offsets land in pre[] (LDS).
template <int K>
__device__ __forceinline__ void chunk_scan(const uint32_t (&v)[K], uint32_t (&ex)[K], uint32_t (&tot)[K],
                                           uint64_t *status, uint32_t maxChunks, uint32_t ch, uint32_t tag,
                                           uint32_t *err, uint32_t *smem, uint32_t *pre) {
  block_excl_scan_multi<1024, K>(v, ex, tot, smem);
  const uint32_t w = threadIdx.x / kWave;
  if (w < (uint32_t)K) {
    uint32_t t = 0;
#pragma unroll
    for (int q = 0; q < K; q++) t = (uint32_t)q == w ? tot[q] : t;
    const uint32_t p = lookback_exclusive_wave(status + (size_t)w * maxChunks, ch, ch, tag, t, false, err);
    if (lane_id() == 0) pre[w] = p;
  }
  __syncthreads();
}

// identity of the child-bound accumulators (min words, then max words)
__global__ void k_cb_init(int32_t *cb, size_t nWords) {
  for (size_t w = blockIdx.x * (size_t)blockDim.x + threadIdx.x; w < nWords; w += (size_t)gridDim.x * blockDim.x)
    cb[w] = (w % 6) < 3 ? enc_f(INFINITY) : enc_f(-INFINITY);
}

__device__ __forceinline__ uint32_t find_seg(const uint32_t *start, uint32_t nSeg, uint32_t t) {
  uint32_t lo = 0, hi = nSeg;  // last s with start[s] <= t
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (start[mid] <= t) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ void seg_cb(const Seg &g, float (&lo)[3], float (&hi)[3]) {
  for (int d = 0; d < 3; d++) { lo[d] = dec_f(g.cb[d]); hi[d] = dec_f(g.cb[d + 3]); }
}

// Bin: one LDS bin set per binning tile, written out as that tile's partial.
__global__ __launch_bounds__(kBlock) void k_bin(Ctx c, int level) {
  __shared__ int32_t sBins[3 * kMaxBins * kBinWords];
  __shared__ uint32_t sSeg;
  const int par = level & 1;
  const uint32_t nbt = c.ctr->numBinTiles;
  const uint32_t nSeg = c.ctr->numLarge[par];
  const Item *src = c.items[par];
  const int nb = c.nb;
  const int W = 3 * nb * kBinWords;
  for (uint32_t bt = blockIdx.x; bt < nbt; bt += gridDim.x) {
    if (threadIdx.x == 0) sSeg = find_seg(c.binStart, nSeg, bt);
    for (int w = threadIdx.x; w < W; w += kBlock) sBins[w] = bin_init(w);
    sync_block_lds();
    const uint32_t s = sSeg;
    const Seg g = c.segs[par][s];
    float clo[3], chi[3];
    seg_cb(g, clo, chi);
    const uint32_t b0 = g.begin + (bt - c.binStart[s]) * kBinTile;
    const uint32_t e0 = min(b0 + (uint32_t)kBinTile, g.end);
    for (uint32_t i = b0 + threadIdx.x; i < e0; i += kBlock) bin_add(sBins, nb, c.binEps, src[i], clo, chi);
    sync_block_lds();
    int32_t *out = c.partial + (size_t)bt * W;
    for (int w = threadIdx.x; w < W; w += kBlock) out[w] = sBins[w];
    sync_block_lds();
  }
}

// Select: reduce a node's partials, evaluate SAH, record the decision and
// the node's box (the union of its bins); leaves are final here.
__global__ __launch_bounds__(kBlock) void k_select(Ctx c, int level) {
  __shared__ int32_t sBins[3 * kMaxBins * kBinWords];
  const int par = level & 1;
  const uint32_t nSeg = c.ctr->numLarge[par];
  const int nb = c.nb;
  const int W = 3 * nb * kBinWords;
  for (uint32_t s = blockIdx.x; s < nSeg; s += gridDim.x) {
    const uint32_t tb = c.binStart[s], te = c.binStart[s + 1];
    for (int w = threadIdx.x; w < W; w += kBlock) {
      const int k = w & (kBinWords - 1);
      int32_t acc = bin_init(w);
      const int32_t *p = c.partial + (size_t)tb * W + w;
      for (uint32_t t = tb; t < te; t++, p += W) {
        const int32_t v = *p;
        acc = k < 3 ? min(acc, v) : (k < 6 ? max(acc, v) : acc + v);
      }
      sBins[w] = acc;
    }
    sync_block_lds();
    if (threadIdx.x < kWave) {
      const Seg g = c.segs[par][s];
      const uint32_t n = g.end - g.begin;
      int bin = 0;
      const int mode = sah_decide_wave(sBins, nb, n, c.cap, c.elh != 0, bin);
      // node box = union of the dimension-0 bins
      float lo[3], hi[3];
      const bool inBin = lane_id() < (uint32_t)nb;
      const int32_t *w = sBins + (inBin ? lane_id() : 0) * kBinWords;
      for (int k = 0; k < 3; k++) {
        lo[k] = dec_f(wave_min_i32(inBin ? w[k] : INT32_MAX));
        hi[k] = dec_f(wave_max_i32(inBin ? w[3 + k] : INT32_MIN));
      }
      if (lane_id() == 0) {
        c.segMode[s] = mode;
        c.segBin[s] = bin;
        if (mode >= 0) {
          c.segLo[s] = dec_f(g.cb[mode]);
          c.segHi[s] = dec_f(g.cb[mode + 3]);
        }
      }
      store_node_wave(&c.nodes[g.node], lo, hi, mode == kModeLeaf ? pack_oc(g.begin, n) : 0ull);
    }
    sync_block_lds();
  }
}

// Partition: ticket-ordered tiles, left side stable, right side reversed,
// child centroid bounds with one atomic per tile and value.
__global__ __launch_bounds__(kBlock) void k_partition(Ctx c, int level) {
  __shared__ uint32_t sr[kItems * (kBlock / kWave) + 1];
  __shared__ float sb[12 * (kBlock / kWave)];
  __shared__ uint32_t sTile, sSeg, sPrefix;
  const int par = level & 1;
  const uint32_t numTiles = c.ctr->numTiles;
  const uint32_t nSeg = c.ctr->numLarge[par];
  const Seg *segs = c.segs[par];
  const Item *src = c.items[par];
  Item *dst = c.items[par ^ 1];
  const uint32_t tag = (uint32_t)level + 1;
  const int nb = c.nb;
  while (true) {
    if (threadIdx.x == 0) {
      // counter cc hands out tiles cc, cc + kTicketCounters, ...
      const uint32_t cc = blockIdx.x % kTicketCounters;
      const uint32_t t = cc + kTicketCounters * atomicAdd(&c.tickets[kTicketStride * cc], 1u);
      sTile = t;
      if (t < numTiles) sSeg = find_seg(c.tileStart, nSeg, t);
    }
    sync_block_lds();
    const uint32_t t = sTile;
    if (t >= numTiles) break;
    const uint32_t s = sSeg;
    const Seg g = segs[s];
    const int mode = c.segMode[s];
    const uint32_t k = t - c.tileStart[s];
    const uint32_t nTilesSeg = c.tileStart[s + 1] - c.tileStart[s];
    const uint32_t tb = g.begin + k * kTile;
    const uint32_t te = min(tb + (uint32_t)kTile, g.end);
    if (mode == kModeLeaf) {
      for (uint32_t i = tb + threadIdx.x; i < te; i += kBlock) c.primIDs[i] = item_id(src[i]);
      sync_block_lds();
      continue;
    }
    const int bin = c.segBin[s];
    const float lo = c.segLo[s], hi = c.segHi[s];
    const uint32_t mid = g.begin + (g.end - g.begin) / 2;

    Item it[kItems];
    bool left[kItems], in[kItems];
#pragma unroll
    for (int j = 0; j < kItems; j++) {
      const uint32_t i = tb + j * kBlock + threadIdx.x;
      in[j] = i < te;
      left[j] = false;
      if (in[j]) {
        it[j] = src[i];
        left[j] = mode == kModeMiddle ? i < mid : sah_bin(item_center(it[j], mode), lo, hi, nb, c.binEps) < bin;
      }
    }
    uint32_t rank[kItems];
    const uint32_t nLeft = block_rank_striped<kBlock, kItems, true>(left, rank, sr);
    if (threadIdx.x < kWave) {
      const uint32_t prefix = lookback_exclusive_wave(c.tileStatus, t, k, tag, nLeft, false, &c.ctr->spinFault);
      if (threadIdx.x == 0) {
        sPrefix = prefix;
        if (k == nTilesSeg - 1) c.segLeft[s] = prefix + nLeft;
      }
    }
    sync_block_lds();
    const uint32_t P = sPrefix;
    const uint32_t rightsBefore = k * kTile - P;
    float v[12];
    for (int j = 0; j < 3; j++) { v[j] = INFINITY; v[j + 3] = -INFINITY; v[j + 6] = INFINITY; v[j + 9] = -INFINITY; }
#pragma unroll
    for (int j = 0; j < kItems; j++) {
      if (!in[j]) continue;
      const uint32_t local = j * kBlock + threadIdx.x;
      const uint32_t d = left[j] ? g.begin + P + rank[j] : g.end - 1 - (rightsBefore + (local - rank[j]));
      dst[d] = it[j];
      const int o = left[j] ? 0 : 6;
      for (int q = 0; q < 3; q++) {
        const float ce = item_center(it[j], q);
        v[o + q] = fminf(v[o + q], ce);
        v[o + 3 + q] = fmaxf(v[o + 3 + q], ce);
      }
    }
    block_reduce_bounds<kBlock, true>(v, sb);
    if (threadIdx.x < 12) {
      const int side = threadIdx.x / 6;
      const int j = threadIdx.x % 6;
      const bool nonEmpty = side == 0 ? (nLeft > 0) : (nLeft < te - tb);
      if (nonEmpty) {
        // (copy t mod the segment's copy count: copies spread a wide
        // segment's atomics, a narrow one uses only as many as its tiles)
        const uint32_t nCopies = min((uint32_t)kCbCopies, nTilesSeg);
        int32_t *dcb = &c.childCb[((size_t)(t % nCopies) * 2 * c.maxLarge + 2 * s + side) * 6 + j];
        if (j < 3) atomicMin(dcb, enc_f(v[threadIdx.x]));
        else atomicMax(dcb, enc_f(v[threadIdx.x]));
      }
    }
    sync_block_lds();
  }
}

// Emit: child ids for split nodes in segment order, next-level large
// segments with their tilings (the next level's plan) and small segments,
// offsets by the chunk scan; the child-bound accumulators it reads are reset
// for the next level.  Every chunk reads the node and small-segment counts
// before it publishes its look-back words, and the last chunk, whose
// look-back waits for all of those, updates them.
__global__ __launch_bounds__(1024) void k_emit(Ctx c, int level) {
  constexpr int K = 6;  // split, large children, small children, tiles, binning tiles, items
  __shared__ uint32_t smem[K * (1024 / kWave + 1)];
  __shared__ uint32_t sPre[K];
  const int par = level & 1;
  const uint32_t n = c.ctr->numLarge[par];
  const uint32_t nChunks = n ? div_up(n, 1024u) : 1u;
  const uint32_t ch = blockIdx.x;
  if (ch >= nChunks) return;
  const uint32_t nodeBase = c.ctr->nodeCount, smallBase = c.ctr->numSmall;
  const Seg *segs = c.segs[par];
  Seg *next = c.segs[par ^ 1];
  const uint32_t s = ch * 1024 + threadIdx.x;
  uint32_t v[K] = {0u, 0u, 0u, 0u, 0u, 0u};
  Seg g;
  uint32_t cb[2], ce[2];
  bool large[2] = {false, false};
  int32_t ccb[2][6];
  if (s < n) {
    g = segs[s];
    if (c.segMode[s] != kModeLeaf) {
      v[0] = 1;
      const uint32_t mid = g.begin + c.segLeft[s];
      cb[0] = g.begin; ce[0] = mid;
      cb[1] = mid; ce[1] = g.end;
      for (int side = 0; side < 2; side++) {
        const uint32_t sz = ce[side] - cb[side];
        large[side] = sz > (uint32_t)kSmall && sz > c.T;
        if (large[side]) {
          v[1]++;
          v[3] += div_up(sz, kTile);
          v[4] += div_up(sz, kBinTile);
          v[5] += sz;
        } else {
          v[2]++;
        }
      }
    }
    // child bounds (the union over the copies the partition used: one per
    // tile up to kCbCopies), and those copies reset to the identity for
    // the next level
    const uint32_t nCopies = min((uint32_t)kCbCopies, div_up(g.end - g.begin, kTile));
    for (int side = 0; side < 2; side++)
      for (int j = 0; j < 6; j++) {
        int32_t w = j < 3 ? enc_f(INFINITY) : enc_f(-INFINITY);
        const int32_t idw = w;
        for (uint32_t cp = 0; cp < nCopies; cp++) {
          int32_t *p = &c.childCb[((size_t)cp * 2 * c.maxLarge + 2 * s + side) * 6 + j];
          const int32_t u = *p;
          w = j < 3 ? min(w, u) : max(w, u);
          *p = idw;
        }
        ccb[side][j] = w;
      }
  }
  uint32_t ex[K], tot[K];
  chunk_scan<K>(v, ex, tot, c.emitStatus, c.maxChunks, ch, (uint32_t)level + 1, &c.ctr->spinFault, smem, sPre);
  if (v[0]) {
    const uint32_t child = nodeBase + 2 * (sPre[0] + ex[0]);
    c.nodes[g.node].oc = pack_oc(child, 0);
    uint32_t li = sPre[1] + ex[1], si = sPre[2] + ex[2];
    uint32_t ti = sPre[3] + ex[3], bi = sPre[4] + ex[4];
    for (int side = 0; side < 2; side++) {
      if (large[side]) {
        Seg ns;
        ns.begin = cb[side]; ns.end = ce[side]; ns.node = child + side; ns.pad = 0;
        ns.pad2[0] = ns.pad2[1] = 0;
        for (int j = 0; j < 6; j++) ns.cb[j] = ccb[side][j];
        next[li] = ns;
        c.tileStart[li] = ti;
        c.binStart[li] = bi;
        ti += div_up(ce[side] - cb[side], kTile);
        bi += div_up(ce[side] - cb[side], kBinTile);
        li++;
      } else {
        c.small[smallBase + si] = SmallSeg{cb[side], ce[side], child + side, (uint32_t)(par ^ 1)};
        si++;
      }
    }
  }
  if (ch == nChunks - 1 && threadIdx.x == 0) {
    c.ctr->nodeCount = nodeBase + 2 * (sPre[0] + tot[0]);
    c.ctr->numSmall = smallBase + sPre[2] + tot[2];
    const uint32_t nNext = sPre[1] + tot[1];
    c.ctr->numLarge[par ^ 1] = nNext;
    c.ctr->numLarge[par] = 0;
    level_plan_totals(c, (uint32_t)level + 1, nNext, sPre[3] + tot[3], sPre[4] + tot[4], sPre[5] + tot[5]);
  }
}

// ------------------------------------------------------------ small nodes
// One wave finishes a node of <= kSmall prims (or a leaf of any size) in
// LDS with the same rules; nodes are numbered locally (local 0 = the node
// itself) into loc[2*begin + j] and renumbered by k_small_scan/k_small_emit.
__device__ void small_segment(const Ctx &c, const SmallSeg &sg, uint32_t segIdx, Item *sIt,
                              int32_t *sBins, uint32_t *sFr, uint32_t *sMe) {
  const uint32_t lane = lane_id();
  const Item *gItems = c.items[sg.buf];
  const uint32_t gb = sg.begin;
  const uint32_t m = sg.end - sg.begin;
  Node *loc = c.loc + 2 * (size_t)gb;
  const int nb = c.nb;
  uint32_t nLocal = 1;
  if (m <= c.T) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = lane; i < m; i += kWave) {
      const Item v = gItems[gb + i];
      c.primIDs[gb + i] = item_id(v);
      lo[0] = fminf(lo[0], v.a.x); lo[1] = fminf(lo[1], v.a.y); lo[2] = fminf(lo[2], v.a.z);
      hi[0] = fmaxf(hi[0], v.b.x); hi[1] = fmaxf(hi[1], v.b.y); hi[2] = fmaxf(hi[2], v.b.z);
    }
    for (int d = 0; d < 3; d++) { lo[d] = wave_min_dpp(lo[d]); hi[d] = wave_max_dpp(hi[d]); }
    store_node_wave(loc, lo, hi, pack_oc(gb, m));
    if (lane == 0) c.smallCount[segIdx] = nLocal;
    return;
  }
  for (uint32_t i = lane; i < m; i += kWave) sIt[i] = gItems[gb + i];
  int sp = 1;
  if (lane == 0) { sFr[0] = m << 16; sMe[0] = 0; }
  sync_wave_lds();
  while (sp > 0) {
    --sp;
    const uint32_t f = sFr[sp];
    const uint32_t me = sMe[sp];
    const uint32_t b = f & 0xffffu, e = f >> 16;
    const uint32_t n = e - b;
    // bounds on enc_cmp integers (fused integer DPP reductions)
    int32_t ble[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, bhe[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    int32_t cle[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, che[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    for (uint32_t i = b + lane; i < e; i += kWave) {
      const Item v = sIt[i];
      ble[0] = min(ble[0], enc_cmp(v.a.x)); ble[1] = min(ble[1], enc_cmp(v.a.y)); ble[2] = min(ble[2], enc_cmp(v.a.z));
      bhe[0] = max(bhe[0], enc_cmp(v.b.x)); bhe[1] = max(bhe[1], enc_cmp(v.b.y)); bhe[2] = max(bhe[2], enc_cmp(v.b.z));
      for (int d = 0; d < 3; d++) {
        const int32_t ce = enc_cmp(item_center(v, d));
        cle[d] = min(cle[d], ce);
        che[d] = max(che[d], ce);
      }
    }
    float blo[3], bhi[3], clo[3], chi[3];
    for (int d = 0; d < 3; d++) {
      blo[d] = dec_f(wave_min_i32(ble[d])); bhi[d] = dec_f(wave_max_i32(bhe[d]));
      clo[d] = dec_f(wave_min_i32(cle[d])); chi[d] = dec_f(wave_max_i32(che[d]));
    }
    int mode = kModeLeaf, bin = 0;
    if (n > c.T) {
      for (int w = lane; w < 3 * nb * kBinWords; w += kWave) sBins[w] = bin_init(w);
      sync_wave_lds();
      for (uint32_t i = b + lane; i < e; i += kWave) bin_add(sBins, nb, c.binEps, sIt[i], clo, chi);
      sync_wave_lds();
      mode = sah_decide_wave(sBins, nb, n, c.cap, c.elh != 0, bin);
    }
    if (mode == kModeLeaf) {
      for (uint32_t i = b + lane; i < e; i += kWave) c.primIDs[gb + i] = item_id(sIt[i]);
      store_node_wave(loc + me, blo, bhi, pack_oc(gb + b, n));
      sync_wave_lds();
      continue;
    }
    const float lo = mode >= 0 ? clo[mode] : 0.f, hi = mode >= 0 ? chi[mode] : 0.f;
    const uint32_t midPos = b + n / 2;
    // partition in place: the node's items (<= kSmall/64 per lane) are read
    // into registers first; a wave's LDS accesses run in program order, so
    // the writes below cannot overtake those reads
    constexpr int kPer = kSmall / kWave;
    Item v[kPer];
    uint32_t dst[kPer];
    bool act[kPer];
    uint32_t nL = 0, nR = 0;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const uint32_t i = b + (uint32_t)k * kWave + lane;
      act[k] = i < e;
      bool isL = false;
      if (act[k]) {
        v[k] = sIt[i];
        isL = mode == kModeMiddle ? i < midPos : sah_bin(item_center(v[k], mode), lo, hi, nb, c.binEps) < bin;
      }
      const uint64_t mL = __ballot(isL);
      const uint64_t mR = __ballot(act[k] && !isL);
      dst[k] = isL ? b + nL + mbcnt(mL) : e - 1 - (nR + mbcnt(mR));
      nL += (uint32_t)__popcll(mL);
      nR += (uint32_t)__popcll(mR);
    }
    sync_wave_lds();
#pragma unroll
    for (int k = 0; k < kPer; k++)
      if (act[k]) sIt[dst[k]] = v[k];
    sync_wave_lds();
    const uint32_t mid = b + nL;
    const uint32_t child = nLocal;
    nLocal += 2;
    store_node_wave(loc + me, blo, bhi, pack_oc(child, 0));
    if (lane == 0) {
      sFr[sp] = mid | (e << 16);
      sMe[sp] = child + 1;
      sFr[sp + 1] = b | (mid << 16);
      sMe[sp + 1] = child;
    }
    sp += 2;
    sync_wave_lds();
  }
  if (lane == 0) c.smallCount[segIdx] = nLocal;
}

// LDS (dynamic, sized by the bin count): items | bins | two stacks
__host__ __device__ constexpr size_t small_lds_bytes(int nb) {
  return sizeof(Item) * kSmall + sizeof(int32_t) * 3 * (size_t)nb * kBinWords + 2 * sizeof(uint32_t) * (kSmall + 2);
}

__global__ __launch_bounds__(kWave) void k_small(Ctx c) {
  extern __shared__ __align__(16) unsigned char smem[];
  Item *sIt = reinterpret_cast<Item *>(smem);
  int32_t *sBins = reinterpret_cast<int32_t *>(smem + sizeof(Item) * kSmall);
  uint32_t *sFr = reinterpret_cast<uint32_t *>(sBins + 3 * c.nb * kBinWords);
  uint32_t *sMe = sFr + (kSmall + 2);
  const uint32_t nSmall = c.ctr->numSmall;
  for (uint32_t w = blockIdx.x; w < nSmall; w += gridDim.x) {
    small_segment(c, c.small[w], w, sIt, sBins, sFr, sMe);
    sync_wave_lds();
  }
}

// global ids of the small segments' local nodes (the final node count goes
// to numNodesOut: every workgroup reads nodeCount as its base)
__global__ __launch_bounds__(1024) void k_small_scan(Ctx c, uint32_t tag) {
  __shared__ uint32_t ss[1024 / kWave + 2];
  count_scan_tile(c.smallCount, c.smallBase, c.ctr->numSmall, c.ctr->nodeCount, c.tileStatus, tag,
                  &c.ctr->numNodesOut, ss, 1u, &c.ctr->spinFault);
}

__global__ __launch_bounds__(kWave) void k_small_emit(Ctx c, Node *out) {
  const uint32_t nSmall = c.ctr->numSmall;
  for (uint32_t w = blockIdx.x; w < nSmall; w += gridDim.x) {
    const SmallSeg sg = c.small[w];
    copy_local_nodes(c.loc + 2 * (size_t)sg.begin, c.smallCount[w], sg.node, c.smallBase[w], out);
  }
}

__global__ void k_copy_top(const Node *src, Node *dst, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace sah

static_assert(sizeof(sah::Counters) <= 64 * sizeof(uint32_t), "host counter mirror");
static uint32_t *sah_host_words() {
  static thread_local uint32_t *p = nullptr;
  if (!p && hipHostMalloc((void **)&p, 64 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess)
    p = nullptr;
  return p;
}

int sah_build(cubql_bvh3f *bvh, const cubql_box3f *boxes, uint32_t n, uint32_t T, uint32_t cap,
              int numBins, hipStream_t s, const cubql_mem_resource *mem, bool elh) {
  using namespace sah;
  if (numBins < 2 || numBins > kMaxBins) {
    set_error("sahBuilder: numBins must be in [2, %d] (got %d)", kMaxBins, numBins);
    return CUBQL_ERR_INVALID_ARG;
  }
  ScratchArena scr(mem, s);
  Ctx c{};
  c.boxes = boxes;
  c.n = n;
  c.cap = std::max<uint32_t>(1u, std::min<uint32_t>(cap, 65535u));
  c.T = std::min(std::max<uint32_t>(T, 1u), c.cap);
  c.nb = numBins;
  c.binEps = elh ? 0.f : 1e-20f;
  c.elh = elh ? 1 : 0;
  const uint32_t numBlocks = div_up(n, kTile);
  c.maxLarge = n / (kSmall + 1) + 2;
  const uint32_t maxTiles = div_up(n, kTile) + c.maxLarge + 1;
  const uint32_t maxBinTiles = div_up(n, kBinTile) + c.maxLarge + 1;
  const size_t binW = 3 * (size_t)numBins * kBinWords;
  const uint32_t maxLevels = 4096;
  CUBQL_CHECK(scr.reserve(&c.items[0], n));
  CUBQL_CHECK(scr.reserve(&c.items[1], n));
  CUBQL_CHECK(scr.reserve(&c.nodes, 2 * (size_t)n + 2));
  CUBQL_CHECK(scr.reserve(&c.ctr, 1));
  CUBQL_CHECK(scr.reserve(&c.blockCount, numBlocks + 1));
  CUBQL_CHECK(scr.reserve(&c.blockBounds, 8 * (size_t)numBlocks + 8));
  CUBQL_CHECK(scr.reserve(&c.segs[0], c.maxLarge));
  CUBQL_CHECK(scr.reserve(&c.segs[1], c.maxLarge));
  CUBQL_CHECK(scr.reserve(&c.childCb, 12 * (size_t)c.maxLarge * kCbCopies));
  CUBQL_CHECK(scr.reserve(&c.tickets, kTicketStride * kTicketCounters));
  CUBQL_CHECK(scr.reserve(&c.segLeft, c.maxLarge));
  CUBQL_CHECK(scr.reserve(&c.tileStart, c.maxLarge + 1));
  CUBQL_CHECK(scr.reserve(&c.binStart, c.maxLarge + 1));
  CUBQL_CHECK(scr.reserve(&c.segMode, c.maxLarge));
  CUBQL_CHECK(scr.reserve(&c.segBin, c.maxLarge));
  CUBQL_CHECK(scr.reserve(&c.segLo, c.maxLarge));
  CUBQL_CHECK(scr.reserve(&c.segHi, c.maxLarge));
  CUBQL_CHECK(scr.reserve(&c.tileStatus, maxTiles));
  CUBQL_CHECK(scr.reserve(&c.partial, (size_t)maxBinTiles * binW));
  CUBQL_CHECK(scr.reserve(&c.small, (size_t)n + 1));
  CUBQL_CHECK(scr.reserve(&c.loc, 2 * (size_t)n + 64));
  CUBQL_CHECK(scr.reserve(&c.smallCount, (size_t)n + 1));
  CUBQL_CHECK(scr.reserve(&c.smallBase, (size_t)n + 1));
  CUBQL_CHECK(scr.reserve(&c.levelItems, maxLevels + 2));
  c.maxChunks = div_up(c.maxLarge, 1024u);
  CUBQL_CHECK(scr.reserve(&c.emitStatus, 6 * (size_t)c.maxChunks));
  CUBQL_CHECK(scr.commit());

  uint32_t *primIDs = nullptr;
  CUBQL_CHECK(mem_alloc(mem, (void **)&primIDs, std::max<size_t>(n, 1) * sizeof(uint32_t), s));
  c.primIDs = primIDs;
  auto fail = [&](int st) {
    mem_free(mem, primIDs, s);
    return st;
  };
  int st = fill_u32((uint32_t *)c.tileStatus, 0u, 2 * (size_t)maxTiles, s);
  if (!st) st = fill_u32((uint32_t *)c.emitStatus, 0u, 12 * (size_t)c.maxChunks, s);
  if (st) return fail(st);
  k_cb_init<<<256, 256, 0, s>>>(c.childCb, 12 * (size_t)c.maxLarge * kCbCopies);
  if (n) {
    k_count<<<numBlocks, kBlock, 0, s>>>(c);
    k_count_scan<<<1, 1024, 0, s>>>(c, numBlocks);
    k_write_items<<<numBlocks, kBlock, 0, s>>>(c);
  } else {
    st = fill_u32((uint32_t *)c.ctr, 0u, sizeof(Counters) / 4, s);
    if (st) return fail(st);
  }
  if (hipGetLastError() != hipSuccess) return fail((set_error("sah init launch failed"), CUBQL_ERR_HIP));

  uint32_t *hw = sah_host_words();
  if (!hw) return fail((set_error("hipHostMalloc failed"), CUBQL_ERR_HIP));
  int dev = 0, numCU = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&numCU, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    numCU = 256;
  // at least one workgroup per ticket counter, so every counter's tiles run
  const uint32_t partGrid = std::max<uint32_t>(kTicketCounters, std::min<uint32_t>(maxTiles, (uint32_t)numCU * 8));
  const uint32_t binGrid = std::min<uint32_t>(maxBinTiles, (uint32_t)numCU * 8);
  const uint32_t selGrid = std::min<uint32_t>(c.maxLarge, (uint32_t)numCU * 8);
  int level = 0;
  std::vector<std::pair<int, int>> binSlots, partSlots;
  bool synced = false;  // hw holds the counters after the last level batch
  while (n) {
    for (int i = 0; i < kLevelsPerSync && level < (int)maxLevels; i++, level++) {
      int pb = prof_begin(s);
      k_bin<<<binGrid, kBlock, 0, s>>>(c, level);
      prof_end(pb, s, "sah_bin");
      if (pb >= 0) binSlots.push_back({pb, level});
      k_select<<<selGrid, kBlock, 0, s>>>(c, level);
      int pp = prof_begin(s);
      k_partition<<<partGrid, kBlock, 0, s>>>(c, level);
      prof_end(pp, s, "sah_partition");
      if (pp >= 0) partSlots.push_back({pp, level});
      k_emit<<<c.maxChunks, 1024, 0, s>>>(c, level);
    }
    if (hipGetLastError() != hipSuccess) return fail((set_error("sah level launch failed"), CUBQL_ERR_HIP));
    if (hipMemcpyAsync(hw, c.ctr, sizeof(Counters), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return fail((set_error("sah level sync failed"), CUBQL_ERR_HIP));
    const Counters *hc = (const Counters *)hw;
    if (hc->spinFault) return fail((set_error("sah build: a look-back wait exceeded its bound"), CUBQL_ERR_INTERNAL));
    synced = true;
    if (hc->numLarge[0] == 0 && hc->numLarge[1] == 0) break;
    if (level >= (int)maxLevels) return fail((set_error("SAH tree deeper than %u levels", maxLevels), CUBQL_ERR_INTERNAL));
  }
  // (after the level loop the counters are already on the host)
  if (!synced && (hipMemcpyAsync(hw, c.ctr, sizeof(Counters), hipMemcpyDeviceToHost, s) != hipSuccess ||
                  hipStreamSynchronize(s) != hipSuccess))
    return fail((set_error("sah sync failed"), CUBQL_ERR_HIP));
  const Counters hc0 = *(const Counters *)hw;
  const uint32_t numSmall = hc0.numSmall, numValid = hc0.numValid;
  const uint32_t topCount = hc0.nodeCount;
  // one workgroup per segment: the hardware dispatcher balances the uneven
  // segment costs far better than a fixed round-robin over resident waves
  const uint32_t smallGrid = std::max<uint32_t>(1u, numSmall);
  int smallSlot = -1;
  if (numSmall) {
    smallSlot = prof_begin(s);
    k_small<<<smallGrid, kWave, small_lds_bytes(numBins), s>>>(c);
    prof_end(smallSlot, s, "sah_small");
    // (status tag beyond every level's)
    k_small_scan<<<div_up(numSmall, kCountScanTile), 1024, 0, s>>>(c, maxLevels + 2);
  }
  if (hipGetLastError() != hipSuccess) return fail((set_error("sah small launch failed"), CUBQL_ERR_HIP));
  if (hipMemcpyAsync(hw, c.ctr, sizeof(Counters), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail((set_error("sah final sync failed"), CUBQL_ERR_HIP));
  if (((const Counters *)hw)->spinFault)
    return fail((set_error("sah build: a look-back wait exceeded its bound"), CUBQL_ERR_INTERNAL));
  const uint32_t numNodes = numValid ? (numSmall ? ((const Counters *)hw)->numNodesOut : topCount) : 0;
  if (prof_on()) {
    // algorithmic bytes: binning reads one 32-byte item per prim of the
    // level; the partition reads and writes it; the small phase reads each
    // item once and writes its primID and its nodes
    std::vector<uint32_t> li(level + 1, 0);
    if (level > 0 && hipMemcpy(li.data(), c.levelItems, sizeof(uint32_t) * level, hipMemcpyDeviceToHost) == hipSuccess) {
      for (auto &p : binSlots) prof_set(p.first, 32.0 * li[p.second], li[p.second]);
      for (auto &p : partSlots) prof_set(p.first, 64.0 * li[p.second], li[p.second]);
    }
    prof_set(smallSlot, 36.0 * numValid + 32.0 * ((double)numNodes - topCount + numSmall), numValid);
    prof_flush();
  }
  Node *nodes = nullptr;
  st = mem_alloc(mem, (void **)&nodes, std::max<size_t>(numNodes, 2) * sizeof(Node), s);
  if (st) return fail(st);
  if (numNodes) {
    k_copy_top<<<std::max<uint32_t>(1u, std::min<uint32_t>(div_up(topCount, 256), 1024)), 256, 0, s>>>(
        c.nodes, nodes, topCount);
    if (numSmall) k_small_emit<<<smallGrid, kWave, 0, s>>>(c, nodes);
    if (hipGetLastError() != hipSuccess) {
      mem_free(mem, nodes, s);
      return fail((set_error("sah emit launch failed"), CUBQL_ERR_HIP));
    }
  }
  bvh->nodes = (cubql_node3f *)nodes;
  bvh->numNodes = numNodes;
  bvh->primIDs = primIDs;
  bvh->numPrims = n;
  return CUBQL_OK;
}

}  // namespace cubql
