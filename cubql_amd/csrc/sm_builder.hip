// sm_builder.hip -- adaptive spatial-median BVH builder for gfx950.
//
// Replaces cuBQL::gpuBuilder's spatial-median path (builder/cuda/
// sm_builder.h:469-614, gpu_builder.h:31-67).  Split rules are those shared
// by the reference GPU and CPU builders (sm_builder.h:185-220, 361-363;
// cpu/spatialMedian.h:81-107): a node with count <= T is a leaf; otherwise
// split the widest centroid-bounds axis (ties to the lower axis) at the
// midpoint; a prim goes right iff centre[dim] >= pos.  Degenerate splits
// follow the CPU builder exactly (one empty side -> split the node's range
// in the middle), so the GPU tree equals cpu::spatialMedian's tree node for
// node, up to node numbering.
//
// MI355X design (not the reference's prim-state + global sort scheme):
//   * prims are kept physically partitioned: each open node owns a
//     contiguous range of 16-byte items {cx,cy,cz,primID} in HBM;
//   * one "partition" pass per tree level moves every prim of a large node
//     to its child range (left stable, right reversed -- the CPU order),
//     using a single-pass decoupled look-back scan over 2048-item tiles and
//     accumulating child centroid bounds with one atomic per tile;
//   * nodes of <= kSmall prims are finished by one wavefront each, entirely
//     in LDS, which also computes all their boxes bottom-up;
//   * the level loop runs on the device: kernels read their work counts from
//     device memory, the host only checks for completion every few levels.
#include "block_ops.h"
#include "common.h"
#include "wave_ops.h"
#include "build_common.h"
#include "profile.h"

#include <vector>

namespace cubql {
namespace sm {

constexpr int kBlock = 256;
constexpr int kItems = 8;
constexpr int kTile = kBlock * kItems;  // 2048 prims per partition tile
constexpr int kSmall = 256;             // max prims of a node a wave finishes in LDS
// Segments of up to kMedium prims leave the level loop too: their wave
// splits them in global memory (between the two item buffers) until the
// pieces fit kSmall.
constexpr int kMedium = 1024;
constexpr int kLevelsPerSync = 8;
// Tiles are claimed from kTicketCounters interleaved counters (counter c
// hands out tiles c, c+kTicketCounters, ...) and child centroid bounds are
// accumulated into kCbCopies copies: both spread same-address atomics, which
// serialise at one memory-side atomic unit per address.
constexpr uint32_t kTicketCounters = 8;
constexpr uint32_t kTicketStride = 32;  // words between counters
constexpr uint32_t kCbCopies = 8;

struct Counters {
  uint32_t numValid;
  int32_t rootCb[6];  // encoded centroid bounds of all valid prims
  uint32_t allPoints; // every valid box is a point equal to its centre
  uint32_t nodeCount;
  uint32_t numNodesOut;   // node count after the small phase
  uint32_t numLarge[2];
  uint32_t numSmall;
  uint32_t ticket;
  uint32_t numTiles;
  uint32_t numTop;    // entries in the top-node list
  uint32_t baseNode[2], baseSmall[2], baseTop[2];  // each level's bases, by level parity (emit)
};

struct Seg {          // an open node with > kMedium prims
  uint32_t begin, end, node, pad;
  int32_t cb[6];      // encoded centroid bounds
  uint32_t pad2[2];
};
static_assert(sizeof(Seg) == 48, "Seg");

struct SmallSeg {     // a node finished by one wave
  uint32_t begin, end, node, buf;
};

struct Ctx {
  uint32_t *tickets;       // kTicketCounters counters, kTicketStride words apart
  const cubql_box3f *boxes;
  uint32_t n;              // number of input boxes
  uint32_t T;              // leaf threshold
  float4 *items[2];
  uint32_t *primIDs;       // output primIDs (n entries)
  Node *nodes;             // temp nodes (capacity 2*numValid+2)
  Counters *ctr;
  uint32_t *blockCount;    // init pass
  float *blockBounds;      // init pass: per-workgroup centroid bounds + point flag (8 floats)
  Node *loc;               // small phase: local nodes of segment s at loc[2*begin_s + j]
  uint32_t *smallCount;    // small phase: local node count per small segment
  uint32_t *smallBase;     // small phase: first global id of local node 1
  // per-level arrays, double-buffered by level parity: the emit of level l
  // reads the [l&1] copies and prepares the [(l+1)&1] ones
  Seg *segs[2];            // large segments
  int32_t *childCb[2];     // [kCbCopies][2*maxLarge][6] child centroid bounds
  uint32_t *segLeft[2];    // [maxLarge] left-side count after the partition
  uint32_t *tileStart[2];  // [maxLarge+1] first partition tile of each segment
  uint32_t *tileSeg[2];    // [maxTiles] segment of each partition tile
  int32_t *segDim[2];      // [maxLarge] split plane
  float *segPos[2];
  uint64_t *emitStatus;    // [4][maxEmitBlocks] emit look-back status per count
  uint32_t maxEmitBlocks;
  uint64_t *tileStatus;    // [maxTiles]
  SmallSeg *small;         // [numValid]
  uint32_t *topNodes;      // [numValid] inner nodes of the level phase
  uint32_t *levelStart;    // [maxLevels+1] start of each level in topNodes
  uint32_t *levelItems;    // [maxLevels] prims moved by each level's partition
  uint32_t maxLarge;
};

// ---------------------------------------------------------------- init pass
// Pass 1: per-tile valid count, root centroid bounds, point-data flag.
// It also writes every prim's item in place, which is the compacted item
// array whenever all boxes are valid; otherwise pass 3 rewrites it.
__global__ __launch_bounds__(kBlock) void k_count(Ctx c) {
  __shared__ float sb[12 * (kBlock / kWave)];
  __shared__ uint32_t ss[kBlock / kWave + 1];
  const uint32_t base = blockIdx.x * kTile;
  uint32_t cnt = 0;
  bool pts = true;
  float v[12];
  for (int j = 0; j < 3; j++) { v[j] = INFINITY; v[j + 3] = -INFINITY; v[j + 6] = INFINITY; v[j + 9] = -INFINITY; }
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    uint32_t i = base + k * kBlock + threadIdx.x;
    if (i < c.n) {
      cubql_box3f b = load_box(c.boxes, i);
      float ce[3];
      for (int d = 0; d < 3; d++) ce[d] = (b.lower[d] + b.upper[d]) / 2.f;
      c.items[0][i] = make_float4(ce[0], ce[1], ce[2], __uint_as_float(i));
      if (box_is_valid(b)) {
        cnt++;
        for (int d = 0; d < 3; d++) {
          v[d] = fminf(v[d], ce[d]);
          v[d + 3] = fmaxf(v[d + 3], ce[d]);
          pts = pts && (b.lower[d] == b.upper[d]) && (ce[d] == b.lower[d]);
        }
      }
    }
  }
  block_reduce_bounds<kBlock>(v, sb);
  uint32_t tot;
  block_excl_scan<kBlock>(cnt, ss, &tot);
  bool allPts = __syncthreads_and(pts);
  // per-workgroup partials, reduced by k_count_scan (no same-address atomics)
  if (threadIdx.x < 6) c.blockBounds[blockIdx.x * 8 + threadIdx.x] = v[threadIdx.x];
  if (threadIdx.x == 0) {
    c.blockCount[blockIdx.x] = tot;
    c.blockBounds[blockIdx.x * 8 + 6] = allPts ? 1.f : 0.f;
  }
}

// Pass 2 (single workgroup): exclusive scan of tile counts, root setup.
__global__ __launch_bounds__(1024) void k_count_scan(Ctx c, uint32_t numBlocks) {
  __shared__ uint32_t ss[1024 / kWave + 1];
  __shared__ float sr[7][1024 / kWave];
  uint32_t run = 0;
  float red[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 1.f};
  for (uint32_t b0 = 0; b0 < numBlocks; b0 += 1024) {
    uint32_t i = b0 + threadIdx.x;
    uint32_t v = i < numBlocks ? c.blockCount[i] : 0;
    if (i < numBlocks) {
      const float *bb = c.blockBounds + (size_t)i * 8;
      for (int j = 0; j < 3; j++) {
        red[j] = fminf(red[j], bb[j]);
        red[j + 3] = fmaxf(red[j + 3], bb[j + 3]);
      }
      red[6] = fminf(red[6], bb[6]);
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan<1024>(v, ss, &tot);
    if (i < numBlocks) c.blockCount[i] = run + ex;
    run += tot;
  }
  for (int j = 0; j < 7; j++) {
    float r = (j >= 3 && j < 6) ? wave_max(red[j]) : wave_min(red[j]);
    if (lane_id() == 0) sr[j][threadIdx.x / kWave] = r;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    Counters *k = c.ctr;
    for (int j = 0; j < 7; j++) {
      float r = sr[j][0];
      for (int w = 1; w < 1024 / kWave; w++) r = (j >= 3 && j < 6) ? fmaxf(r, sr[j][w]) : fminf(r, sr[j][w]);
      if (j < 6) k->rootCb[j] = enc_f(r);
      else k->allPoints = r > 0.5f ? 1u : 0u;
    }
    k->numValid = run;
    k->nodeCount = 2;
    k->numLarge[0] = k->numLarge[1] = 0;
    k->numSmall = 0;
    k->numTop = 0;
    c.levelStart[0] = 0;
    Node z;
    for (int d = 0; d < 3; d++) { z.lo[d] = 0.f; z.hi[d] = 0.f; }
    z.oc = 0;
    c.nodes[1] = z;  // node 1 is never used (bvh.h:66-70)
    k->numTiles = 0;
    k->ticket = 0;
    for (uint32_t cc = 0; cc < kTicketCounters; cc++) c.tickets[kTicketStride * cc] = 0;
    if (run > kMedium && run > c.T) {
      Seg s;
      s.begin = 0; s.end = run; s.node = 0; s.pad = 0; s.pad2[0] = s.pad2[1] = 0;
      for (int j = 0; j < 6; j++) s.cb[j] = k->rootCb[j];
      c.segs[0][0] = s;
      k->numLarge[0] = 1;
      // plan of level 0
      float lo[3], hi[3];
      for (int d = 0; d < 3; d++) { lo[d] = dec_f(s.cb[d]); hi[d] = dec_f(s.cb[d + 3]); }
      int dim;
      float pos;
      choose_split(lo, hi, dim, pos);
      c.segDim[0][0] = dim;
      c.segPos[0][0] = pos;
      c.segLeft[0][0] = 0;
      for (uint32_t cp = 0; cp < kCbCopies; cp++)
        for (int j = 0; j < 12; j++) c.childCb[0][(size_t)cp * 12 * c.maxLarge + j] = enc_f((j % 6) < 3 ? INFINITY : -INFINITY);
      const uint32_t tiles = div_up(run, kTile);
      c.tileStart[0][0] = 0;
      c.tileStart[0][1] = tiles;
      k->numTiles = tiles;
      c.levelItems[0] = run;
    } else if (run > 0) {
      SmallSeg s{0, run, 0, 0};
      c.small[0] = s;
      k->numSmall = 1;
    }
    // level 0's bases for the emit
    k->baseNode[0] = k->nodeCount;
    k->baseSmall[0] = k->numSmall;
    k->baseTop[0] = k->numTop;
  }
  __syncthreads();
  const uint32_t tiles = c.ctr->numTiles;
  for (uint32_t t = threadIdx.x; t < tiles; t += 1024) c.tileSeg[0][t] = 0;
}

// Pass 3: write compacted items (ascending primID) and the invalid tail;
// nothing to do when every box was valid.
__global__ __launch_bounds__(kBlock) void k_write_items(Ctx c) {
  if (c.ctr->numValid == c.n) return;
  __shared__ uint32_t sr[kItems * (kBlock / kWave) + 1];
  const uint32_t base = blockIdx.x * kTile;
  bool valid[kItems];
  float4 it[kItems];
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    uint32_t i = base + k * kBlock + threadIdx.x;
    valid[k] = false;
    if (i < c.n) {
      cubql_box3f b = load_box(c.boxes, i);
      valid[k] = box_is_valid(b);
      it[k] = make_float4((b.lower[0] + b.upper[0]) / 2.f, (b.lower[1] + b.upper[1]) / 2.f,
                          (b.lower[2] + b.upper[2]) / 2.f, __uint_as_float(i));
    }
  }
  uint32_t rank[kItems];
  block_rank_striped<kBlock, kItems>(valid, rank, sr);
  const uint32_t off = c.blockCount[blockIdx.x];
  const uint32_t numValid = c.ctr->numValid;
  const uint32_t invBefore = base - off;
#pragma unroll
  for (int k = 0; k < kItems; k++) {
    uint32_t i = base + k * kBlock + threadIdx.x;
    if (i >= c.n) continue;
    uint32_t local = k * kBlock + threadIdx.x;
    if (valid[k]) c.items[0][off + rank[k]] = it[k];
    else c.primIDs[numValid + invBefore + (local - rank[k])] = i;
  }
}

// ------------------------------------------------------------- level pass
// Partition, software-pipelined over a workgroup's tiles.  A workgroup
// claims tiles in ticket order (so a tile only ever waits on tiles claimed
// by running workgroups); for its current tile it ranks the items left /
// right, stages them in LDS in output order, resolves its prefix by
// decoupled look-back among the tiles of its node and copies the staged
// items out in contiguous runs.  After its look-back the workgroup claims
// and describes its next tile, whose items are loaded into registers
// before the current tile's copy-out, so each tile's load latency overlaps
// the previous tile's stores.  (Claiming only then keeps the claim order
// close to the order in which tiles publish their counts, which shortens
// every later tile's look-back.)
//
// A workgroup holds at most two tiles, the later one claimed after the
// earlier; every wait is on a lower tile, so the lowest unfinished tile
// never waits and the grid drains.
struct TileDesc {   // the next tile, written by one thread, read after a barrier
  uint32_t t, seg, begin, end, dim, pos, tileFirst, tileEnd;
};

__global__ __launch_bounds__(kBlock) void k_partition(Ctx c, int level) {
  __shared__ float4 sStage[kTile];
  __shared__ uint32_t sr[kItems * (kBlock / kWave) + 1];
  __shared__ float sb[12 * (kBlock / kWave)];
  __shared__ TileDesc sNext;
  __shared__ uint32_t sPrefix;
  const int par = level & 1;
  const uint32_t numTiles = c.ctr->numTiles;
  if (numTiles == 0) return;  // a level past the tree's depth
  const float4 *src = c.items[par];
  float4 *dst = c.items[par ^ 1];
  const uint32_t tag = (uint32_t)level + 1;
  // counter cc hands out tiles cc, cc + kTicketCounters, ...: every tile a
  // claimed tile waits on is either claimed (and held by a running
  // workgroup) or next in its counter, which running workgroups serve
  const uint32_t cc = blockIdx.x % kTicketCounters;
  auto fetch = [&]() {
    TileDesc d;
    d.t = cc + kTicketCounters * atomicAdd(&c.tickets[kTicketStride * cc], 1u);
    if (d.t < numTiles) {
      d.seg = c.tileSeg[par][d.t];
      const Seg g = c.segs[par][d.seg];
      d.begin = g.begin;
      d.end = g.end;
      d.dim = (uint32_t)c.segDim[par][d.seg];
      d.pos = __float_as_uint(c.segPos[par][d.seg]);
      d.tileFirst = c.tileStart[par][d.seg];
      d.tileEnd = c.tileStart[par][d.seg + 1];
    }
    sNext = d;
  };
  if (threadIdx.x == 0) fetch();
  sync_block_lds();
  TileDesc cur = sNext;
  if (cur.t >= numTiles) return;
  float4 it[kItems];
  {
    const uint32_t tb = cur.begin + (cur.t - cur.tileFirst) * kTile;
    const uint32_t last = min(tb + (uint32_t)kTile, cur.end) - 1;
#pragma unroll
    for (int j = 0; j < kItems; j++) it[j] = src[min(tb + j * kBlock + threadIdx.x, last)];
  }
  while (true) {
    const uint32_t k = cur.t - cur.tileFirst;
    const uint32_t tb = cur.begin + k * kTile;
    const uint32_t n = min(tb + (uint32_t)kTile, cur.end) - tb;
    const int dim = (int)cur.dim;
    const float pos = __uint_as_float(cur.pos);
    bool left[kItems];
#pragma unroll
    for (int j = 0; j < kItems; j++) left[j] = j * kBlock + threadIdx.x < n && coord(it[j], dim) < pos;
    uint32_t rank[kItems];
    // its barriers also order the previous tile's reads of sStage, sNext
    // and sPrefix before this tile's writes
    const uint32_t nLeft = block_rank_striped<kBlock, kItems, true>(left, rank, sr);

    // stage in output order: lefts ascending, rights reversed (the right
    // side fills its range from the end)
#pragma unroll
    for (int j = 0; j < kItems; j++) {
      const uint32_t l = j * kBlock + threadIdx.x;
      if (l < n) sStage[left[j] ? rank[j] : n - 1 - (l - rank[j])] = it[j];
    }
    if (threadIdx.x < kWave) {
      const uint32_t prefix = lookback_exclusive_wave(c.tileStatus, cur.t, k, tag, nLeft);
      if (threadIdx.x == 0) {
        sPrefix = prefix;
        if (cur.t == cur.tileEnd - 1) c.segLeft[par][cur.seg] = prefix + nLeft;
        fetch();
      }
    }
    sync_block_lds();
    const uint32_t P = sPrefix;
    const TileDesc nxt = sNext;
    // the next tile's loads go out before this tile's stores (the asm
    // statement keeps the compiler from sinking them below the copy-out;
    // lanes past a short tile re-read its last item)
    if (nxt.t < numTiles) {
      const uint32_t tb2 = nxt.begin + (nxt.t - nxt.tileFirst) * kTile;
      const uint32_t last = min(tb2 + (uint32_t)kTile, nxt.end) - 1;
#pragma unroll
      for (int j = 0; j < kItems; j++) it[j] = src[min(tb2 + j * kBlock + threadIdx.x, last)];
      asm volatile("" ::: "memory");
    }
    // copy out: staged [0, nLeft) -> begin + P + i; staged [nLeft, n) ->
    // the right run, which ends rightsBefore items before the node's end;
    // child centroid bounds on the way (left lo/hi, right lo/hi)
    const uint32_t leftBase = cur.begin + P;
    const uint32_t rightBase = cur.end - (k * kTile - P) - n;
    float v[12];
    {
      float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll 4
      for (uint32_t i = threadIdx.x; i < nLeft; i += kBlock) {
        const float4 x = sStage[i];
        dst[leftBase + i] = x;
        lo[0] = fminf(lo[0], x.x); lo[1] = fminf(lo[1], x.y); lo[2] = fminf(lo[2], x.z);
        hi[0] = fmaxf(hi[0], x.x); hi[1] = fmaxf(hi[1], x.y); hi[2] = fmaxf(hi[2], x.z);
      }
      for (int d = 0; d < 3; d++) { v[d] = lo[d]; v[d + 3] = hi[d]; }
    }
    {
      float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll 4
      for (uint32_t i = nLeft + threadIdx.x; i < n; i += kBlock) {
        const float4 x = sStage[i];
        dst[rightBase + i] = x;
        lo[0] = fminf(lo[0], x.x); lo[1] = fminf(lo[1], x.y); lo[2] = fminf(lo[2], x.z);
        hi[0] = fmaxf(hi[0], x.x); hi[1] = fmaxf(hi[1], x.y); hi[2] = fmaxf(hi[2], x.z);
      }
      for (int d = 0; d < 3; d++) { v[d + 6] = lo[d]; v[d + 9] = hi[d]; }
    }
    block_reduce_bounds<kBlock, true>(v, sb);
    if (threadIdx.x < 12) {
      const int side = threadIdx.x / 6;
      const int j = threadIdx.x % 6;
      const bool nonEmpty = side == 0 ? (nLeft > 0) : (nLeft < n);
      if (nonEmpty) {
        int32_t *dstCb =
            &c.childCb[par][((size_t)(blockIdx.x % kCbCopies) * 2 * c.maxLarge + 2 * cur.seg + side) * 6 + j];
        if (j < 3) atomicMin(dstCb, enc_f(v[threadIdx.x]));
        else atomicMax(dstCb, enc_f(v[threadIdx.x]));
      }
    }
    if (nxt.t >= numTiles) break;
    cur = nxt;
  }
}

// Emit, in three passes over the level's segments (count, scan, write), so
// levels with tens of thousands of segments run on the whole chip: node
// records for the split nodes, the next level's segments (with their split
// planes, tile offsets and tile->segment map: the next level's plan), the
// small-segment list and fixups for one-sided (degenerate) splits.  Output
// order is segment order, so numbering is deterministic.
constexpr int kEmitBlock = 256;

struct EmitSeg {
  Seg g;
  uint32_t cb[2], ce[2];
  bool large[2];
  bool oneSided;
};

__device__ __forceinline__ void emit_classify(const Ctx &c, int par, uint32_t s, EmitSeg &e) {
  e.g = c.segs[par][s];
  const uint32_t size = e.g.end - e.g.begin;
  const uint32_t nl = c.segLeft[par][s];
  e.oneSided = (nl == 0 || nl == size);
  const uint32_t mid = e.oneSided ? (uint32_t)(((uint64_t)e.g.begin + e.g.end) / 2) : e.g.begin + nl;
  e.cb[0] = e.g.begin; e.ce[0] = mid;
  e.cb[1] = mid; e.ce[1] = e.g.end;
  for (int side = 0; side < 2; side++) {
    const uint32_t sz = e.ce[side] - e.cb[side];
    e.large[side] = sz > (uint32_t)kMedium && sz > c.T;
  }
}

// packed per-segment counts: large children | small children << 10 | fixups << 20
__device__ __forceinline__ void emit_counts(const EmitSeg &e, uint32_t &lsf, uint32_t &tiles, uint32_t &items) {
  lsf = 0; tiles = 0; items = 0;
  for (int side = 0; side < 2; side++) {
    if (e.large[side]) {
      lsf += 1u + (e.oneSided ? (1u << 20) : 0u);
      tiles += div_up(e.ce[side] - e.cb[side], kTile);
      items += e.ce[side] - e.cb[side];
    } else {
      lsf += 1u << 10;
    }
  }
}

// One emit launch per level: each workgroup classifies kEmitBlock segments,
// scans their counts (next-level large segments, small segments, tiles,
// prims) within the workgroup, gets its offsets from the workgroups before
// it by decoupled look-back (one wave per count; workgroups only wait on
// lower ones, which the dispatcher starts first), and writes the node
// records, the next level's segments and plan, the tile->segment map and
// the small-segment list.  Children of one-sided (degenerate) splits are
// halves of the node's range whose centroid bounds the partition did not
// see: the workgroup reduces them from the prims and plans them itself.
// The last workgroup writes the level totals and the next level's bases
// (into the other parity's slots) and clears this level's segment count;
// every workgroup has read that count before it published, and the last
// one has waited for all of them.  A level with no segments only carries
// its bookkeeping forward.
__global__ __launch_bounds__(kEmitBlock) void k_emit(Ctx c, int level) {
  __shared__ uint32_t ss[kEmitBlock / kWave + 1];
  __shared__ uint32_t sPre[4];
  __shared__ uint3 sFix[2 * kEmitBlock];  // (segment, begin, end) of one-sided halves
  __shared__ uint32_t sNumFix;
  __shared__ float sb[12 * (kEmitBlock / kWave)];
  const int par = level & 1, np = par ^ 1;
  Counters *k = c.ctr;
  // (read once: the last workgroup clears it)
  const uint32_t n = __hip_atomic_load(&k->numLarge[par], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (n == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      k->numLarge[np] = 0;
      k->numTiles = 0;
      k->baseNode[np] = k->baseNode[par];
      k->baseSmall[np] = k->baseSmall[par];
      k->baseTop[np] = k->baseTop[par];
      c.levelStart[level + 1] = k->baseTop[par];
      c.levelItems[level + 1] = 0;
    }
    return;
  }
  if (blockIdx.x * kEmitBlock >= n) return;
  const uint32_t nb = div_up(n, kEmitBlock);
  const uint32_t s = blockIdx.x * kEmitBlock + threadIdx.x;
  const uint32_t nodeBase = k->baseNode[par], smallBase = k->baseSmall[par], topBase = k->baseTop[par];
  const bool pts = k->allPoints != 0;
  const uint32_t tag = (uint32_t)level + 1;
  EmitSeg e;
  uint32_t nL = 0, nS = 0, nT = 0, nI = 0;
  if (s < n) {
    emit_classify(c, par, s, e);
    for (int side = 0; side < 2; side++) {
      if (e.large[side]) {
        nL++;
        nT += div_up(e.ce[side] - e.cb[side], kTile);
        nI += e.ce[side] - e.cb[side];
      } else {
        nS++;
      }
    }
  }
  if (threadIdx.x == 0) sNumFix = 0;
  uint32_t tL, tS, tT, tI;
  const uint32_t exL = block_excl_scan<kEmitBlock, true>(nL, ss, &tL);
  const uint32_t exS = block_excl_scan<kEmitBlock, true>(nS, ss, &tS);
  const uint32_t exT = block_excl_scan<kEmitBlock, true>(nT, ss, &tT);
  const uint32_t exI = block_excl_scan<kEmitBlock, true>(nI, ss, &tI);
  const uint32_t w = threadIdx.x / kWave;
  if (w < 4) {
    const uint32_t v = w == 0 ? tL : (w == 1 ? tS : (w == 2 ? tT : tI));
    const uint32_t pre =
        lookback_exclusive_wave(c.emitStatus + (size_t)w * c.maxEmitBlocks, blockIdx.x, blockIdx.x, tag, v);
    if (lane_id() == 0) sPre[w] = pre;
  }
  sync_block_lds();
  const uint32_t preL = sPre[0], preS = sPre[1], preT = sPre[2], preI = sPre[3];
  if (blockIdx.x == nb - 1 && threadIdx.x == 0) {
    const uint32_t runL = preL + tL, runS = preS + tS, runT = preT + tT, runI = preI + tI;
    k->nodeCount = k->baseNode[np] = nodeBase + 2 * n;
    k->numSmall = k->baseSmall[np] = smallBase + runS;
    k->numTop = k->baseTop[np] = topBase + n;
    k->numLarge[np] = runL;
    k->numLarge[par] = 0;
    c.levelStart[level + 1] = topBase + n;
    k->numTiles = runT;
    k->ticket = 0;
    for (uint32_t cc = 0; cc < kTicketCounters; cc++) c.tickets[kTicketStride * cc] = 0;
    c.tileStart[np][runL] = runT;
    c.levelItems[level + 1] = runI;
  }
  uint32_t fillStart[2] = {0, 0}, fillCount[2] = {0, 0}, fillSeg[2] = {0, 0};
  if (s < n) {
    uint32_t li = preL + exL, si = preS + exS, ti = preT + exT;
    const uint32_t child = nodeBase + 2 * s;
    if (pts) {
      // point input: the node's box is its centroid bounds, final here
      Node nd;
      for (int d = 0; d < 3; d++) { nd.lo[d] = dec_f(e.g.cb[d]); nd.hi[d] = dec_f(e.g.cb[d + 3]); }
      nd.oc = pack_oc(child, 0);
      c.nodes[e.g.node] = nd;
    } else {
      c.nodes[e.g.node].oc = pack_oc(child, 0);
    }
    c.topNodes[topBase + s] = e.g.node;
    Seg *next = c.segs[np];
    for (int side = 0; side < 2; side++) {
      if (e.large[side]) {
        Seg ns;
        ns.begin = e.cb[side]; ns.end = e.ce[side]; ns.node = child + side; ns.pad = 0;
        ns.pad2[0] = ns.pad2[1] = 0;
        for (int j = 0; j < 6; j++) {
          int32_t v = c.childCb[par][(2 * s + side) * 6 + j];
          for (uint32_t cp = 1; cp < kCbCopies; cp++) {
            const int32_t w2 = c.childCb[par][((size_t)cp * 2 * c.maxLarge + 2 * s + side) * 6 + j];
            v = j < 3 ? min(v, w2) : max(v, w2);
          }
          ns.cb[j] = v;
        }
        next[li] = ns;
        // plan of the next level (one-sided children are planned below)
        float lo[3], hi[3];
        for (int d = 0; d < 3; d++) { lo[d] = dec_f(ns.cb[d]); hi[d] = dec_f(ns.cb[d + 3]); }
        int dim;
        float pos;
        choose_split(lo, hi, dim, pos);
        c.segDim[np][li] = dim;
        c.segPos[np][li] = pos;
        c.segLeft[np][li] = 0;
        for (uint32_t cp = 0; cp < kCbCopies; cp++)
          for (int j = 0; j < 12; j++)
            c.childCb[np][(size_t)cp * 12 * c.maxLarge + li * 12 + j] = enc_f((j % 6) < 3 ? INFINITY : -INFINITY);
        const uint32_t tiles = div_up(ns.end - ns.begin, kTile);
        c.tileStart[np][li] = ti;
        fillStart[side] = ti;
        fillCount[side] = tiles;
        fillSeg[side] = li;
        ti += tiles;
        if (e.oneSided) sFix[atomicAdd(&sNumFix, 1u)] = make_uint3(li, ns.begin, ns.end);
        li++;
      } else {
        c.small[smallBase + si] = SmallSeg{e.cb[side], e.ce[side], child + side, (uint32_t)np};
        si++;
      }
    }
  }
  // tile -> segment map: runs by their own lane, except the long runs of
  // the top levels' few big segments, which the whole wave writes
  constexpr uint32_t kOwnFill = 256;
  const uint32_t lane = lane_id();
  for (int side = 0; side < 2; side++) {
    if (fillCount[side] <= kOwnFill)
      for (uint32_t t = 0; t < fillCount[side]; t++) c.tileSeg[np][fillStart[side] + t] = fillSeg[side];
    uint64_t m = __ballot(fillCount[side] > kOwnFill);
    while (m) {
      const int src = __builtin_ctzll(m);
      m &= m - 1;
      const uint32_t st = __shfl(fillStart[side], src, kWave), cnt = __shfl(fillCount[side], src, kWave);
      const uint32_t sg = __shfl(fillSeg[side], src, kWave);
      for (uint32_t t = lane; t < cnt; t += kWave) c.tileSeg[np][st + t] = sg;
    }
  }
  // one-sided splits: centroid bounds and plane of each half from its prims
  // (a full barrier: the segment records above must land before these
  // overwrite their bounds)
  __syncthreads();
  const uint32_t nFix = sNumFix;
  const float4 *items = c.items[np];
  for (uint32_t f = 0; f < nFix; f++) {
    const uint32_t li = sFix[f].x, gb = sFix[f].y, ge = sFix[f].z;
    Seg &g = c.segs[np][li];
    float v[12];
    for (int j = 0; j < 3; j++) { v[j] = INFINITY; v[j + 3] = -INFINITY; v[j + 6] = INFINITY; v[j + 9] = -INFINITY; }
    for (uint32_t i = gb + threadIdx.x; i < ge; i += kEmitBlock) {
      const float4 it = items[i];
      v[0] = fminf(v[0], it.x); v[1] = fminf(v[1], it.y); v[2] = fminf(v[2], it.z);
      v[3] = fmaxf(v[3], it.x); v[4] = fmaxf(v[4], it.y); v[5] = fmaxf(v[5], it.z);
    }
    block_reduce_bounds<kEmitBlock>(v, sb);
    if (threadIdx.x < 6) g.cb[threadIdx.x] = enc_f(v[threadIdx.x]);
    if (threadIdx.x == 0) {
      const float lo[3] = {v[0], v[1], v[2]}, hi[3] = {v[3], v[4], v[5]};
      int dim;
      float pos;
      choose_split(lo, hi, dim, pos);
      c.segDim[np][li] = dim;
      c.segPos[np][li] = pos;
    }
  }
}

// ------------------------------------------------------------ small nodes
// A small segment (<= kSmall prims, or any leaf-sized node) is finished by
// one wavefront.  Its nodes are numbered locally (local 0 = the segment's
// root, whose global id is already fixed) and written to loc[2*begin + j];
// inner nodes hold the LOCAL index of their first child.  k_small_scan /
// k_small_emit then assign contiguous global ids in segment order and copy
// the nodes to the final array, so no global counter is touched per node and
// the node numbering is deterministic.

struct SmallNodeOut { Node *loc; };

// a pending node of the small phase's LDS loop in one word: range [b, e)
// (<= kSmall) and local id (< 2 * kSmall)
static_assert(kSmall <= 256 && 2 * kMedium <= (1 << 14), "small-phase stack packing");
__device__ __forceinline__ uint32_t fr_pack(uint32_t b, uint32_t e, uint32_t me) {
  return b | (e << 9) | (me << 18);
}

// Segmented inclusive min/max scan over contiguous lane segments: lane l
// ends with the reduction over [head(l), l].  Steps are the DPP wave scan
// (row_shr 1,2,4,8 inside rows of 16, then row_bcast 15 and 31 across rows);
// cond[k] says whether step k's source lane lies in the lane's segment.
// (on enc_cmp-encoded values: integer min/max with an identity as the DPP
// "old" value is one fused v_min/v_max_i32_dpp per step, plus the select)
template <bool MIN>
__device__ __forceinline__ int32_t seg_scan(int32_t v, const bool (&cond)[6]) {
  const int32_t I = MIN ? INT32_MAX : INT32_MIN;
#define CUBQL_SEG_STEP(K, CTRL, ROWMASK)                                                     \
  {                                                                                          \
    const int32_t t = __builtin_amdgcn_update_dpp(I, v, CTRL, ROWMASK, 0xF, false);          \
    const int32_t m = MIN ? min(v, t) : max(v, t);                                           \
    v = cond[K] ? m : v;                                                                     \
  }
  CUBQL_SEG_STEP(0, 0x111, 0xF)
  CUBQL_SEG_STEP(1, 0x112, 0xF)
  CUBQL_SEG_STEP(2, 0x114, 0xF)
  CUBQL_SEG_STEP(3, 0x118, 0xF)
  CUBQL_SEG_STEP(4, 0x142, 0xA)
  CUBQL_SEG_STEP(5, 0x143, 0xC)
#undef CUBQL_SEG_STEP
  return v;
}

// Reduction of v over the lane's segment [sb, se), in every lane of it.
template <bool MIN>
__device__ __forceinline__ int32_t seg_reduce(int32_t v, const bool (&cond)[6], uint32_t last) {
  return __shfl(seg_scan<MIN>(v, cond), (int)last, kWave);
}

__device__ __forceinline__ void store_node_lane(Node *dst, const float (&lo)[3], const float (&hi)[3],
                                                uint64_t oc) {
  Node nd;
  for (int d = 0; d < 3; d++) { nd.lo[d] = lo[d]; nd.hi[d] = hi[d]; }
  nd.oc = oc;
  *dst = nd;
}

// Builds the subtree of a node with cnt <= 64 prims held one per lane
// (lane j = prim at position pos0 + j), splitting every open node of the
// subtree in the same round: each lane tracks the range [sb, se) of the node
// that holds its position; bounds are segmented wave reductions, ranks come
// from ballots masked to the segment, items move with one ds_permute per
// value.  Split rules and partition order (left stable, right reversed,
// one-sided -> middle) are those of the CPU recursion
// (cpu/spatialMedian.h:81-114), so positions and leaves are identical to a
// node-at-a-time build; local ids are assigned breadth-first.
template <bool PTS>
__device__ void reg_subtree(const Ctx &c, uint32_t pos0, uint32_t cnt, float x, float y, float z,
                            uint32_t id, float bx0, float by0, float bz0, float bx1, float by1,
                            float bz1, uint32_t rootLocal, uint32_t &nLocal, Node *loc) {
  const uint32_t lane = lane_id();
  const uint64_t lt = lanemask_lt();
  const uint32_t T = c.T;
  bool active = lane < cnt;
  uint32_t sb = active ? 0u : lane, se = active ? cnt : lane + 1u;
  uint32_t me = rootLocal;
  // coordinates as enc_cmp integers: bounds by fused integer DPP scans,
  // and the plane test in the same order as the float compare
  int32_t xe = enc_cmp(x), ye = enc_cmp(y), ze = enc_cmp(z);
  int32_t b0 = 0, b1 = 0, b2 = 0, b3 = 0, b4 = 0, b5 = 0;
  if (!PTS) {
    b0 = enc_cmp(bx0); b1 = enc_cmp(by0); b2 = enc_cmp(bz0);
    b3 = enc_cmp(bx1); b4 = enc_cmp(by1); b5 = enc_cmp(bz1);
  }
  while (__ballot(active)) {
    const int h = (int)sb, l = (int)lane;
    const bool cond[6] = {l - 1 >= h, l - 2 >= h, l - 4 >= h, l - 8 >= h,
                          (int)(lane & ~15u) - 1 >= h, lane >= 32u && 31 >= h};
    const uint32_t last = se - 1;
    float clo[3], chi[3], nlo[3], nhi[3];
    clo[0] = dec_f(seg_reduce<true>(xe, cond, last));
    clo[1] = dec_f(seg_reduce<true>(ye, cond, last));
    clo[2] = dec_f(seg_reduce<true>(ze, cond, last));
    chi[0] = dec_f(seg_reduce<false>(xe, cond, last));
    chi[1] = dec_f(seg_reduce<false>(ye, cond, last));
    chi[2] = dec_f(seg_reduce<false>(ze, cond, last));
    if (PTS) {
      for (int d = 0; d < 3; d++) { nlo[d] = clo[d]; nhi[d] = chi[d]; }
    } else {
      nlo[0] = dec_f(seg_reduce<true>(b0, cond, last));
      nlo[1] = dec_f(seg_reduce<true>(b1, cond, last));
      nlo[2] = dec_f(seg_reduce<true>(b2, cond, last));
      nhi[0] = dec_f(seg_reduce<false>(b3, cond, last));
      nhi[1] = dec_f(seg_reduce<false>(b4, cond, last));
      nhi[2] = dec_f(seg_reduce<false>(b5, cond, last));
    }
    const uint32_t n = se - sb;
    const bool isHead = lane == sb;
    const bool leaf = active && n <= T;
    const bool split = active && !leaf;
    if (leaf) {
      c.primIDs[pos0 + lane] = id;
      if (isHead) store_node_lane(loc + me, nlo, nhi, pack_oc(pos0 + sb, n));
    }
    int dim;
    float pos;
    choose_split(clo, chi, dim, pos);
    const int32_t ve = dim == 0 ? xe : (dim == 1 ? ye : ze);
    const bool isL = split && ve < enc_cmp(pos);
    const uint64_t segMask = (se >= 64u ? ~0ull : ((1ull << se) - 1ull)) & ~((1ull << sb) - 1ull);
    const uint64_t mL = __ballot(isL) & segMask;
    const uint64_t mR = __ballot(split && !isL) & segMask;
    const uint32_t nl = (uint32_t)__popcll(mL);
    const uint32_t dst = !split ? lane
                                : (isL ? sb + (uint32_t)__popcll(mL & lt) : se - 1u - (uint32_t)__popcll(mR & lt));
    const uint32_t mid = (nl != 0 && nl != n) ? sb + nl
                                              : (uint32_t)(((uint64_t)pos0 + sb + pos0 + se) / 2 - pos0);  // spatialMedian.h:107
    const uint64_t heads = __ballot(split && isHead);
    const uint32_t child = (uint32_t)__shfl((int)(nLocal + 2u * (uint32_t)__popcll(heads & lt)), (int)sb, kWave);
    nLocal += 2u * (uint32_t)__popcll(heads);
    if (split && isHead) store_node_lane(loc + me, nlo, nhi, pack_oc(child, 0));
    xe = (int32_t)lane_push_u(dst, (uint32_t)xe);
    ye = (int32_t)lane_push_u(dst, (uint32_t)ye);
    ze = (int32_t)lane_push_u(dst, (uint32_t)ze);
    id = lane_push_u(dst, id);
    if (!PTS) {
      b0 = (int32_t)lane_push_u(dst, (uint32_t)b0); b1 = (int32_t)lane_push_u(dst, (uint32_t)b1);
      b2 = (int32_t)lane_push_u(dst, (uint32_t)b2); b3 = (int32_t)lane_push_u(dst, (uint32_t)b3);
      b4 = (int32_t)lane_push_u(dst, (uint32_t)b4); b5 = (int32_t)lane_push_u(dst, (uint32_t)b5);
    }
    if (split) {
      if (lane < mid) { se = mid; me = child; }
      else { sb = mid; me = child + 1u; }
    }
    if (leaf) active = false;
  }
}

// Finishes the node of items [gb, gb + m) of gItems (m <= kSmall, or a leaf
// of any size up to T) whose local id is rootMe, numbering new nodes from
// nLocal into loc.
template <bool PTS>
__device__ void small_sub(const Ctx &c, const float4 *gItems, uint32_t gb, uint32_t m, uint32_t rootMe,
                          uint32_t &nLocal, Node *loc, float4 *sIt, uint32_t *sFr) {
  const uint32_t lane = lane_id();
  const uint32_t T = c.T;
  if (m <= T && m > 64) {
    // leaf of any size up to T
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = lane; i < m; i += kWave) {
      float4 v = gItems[gb + i];
      c.primIDs[gb + i] = __float_as_uint(v.w);
      if (PTS) {
        lo[0] = fminf(lo[0], v.x); lo[1] = fminf(lo[1], v.y); lo[2] = fminf(lo[2], v.z);
        hi[0] = fmaxf(hi[0], v.x); hi[1] = fmaxf(hi[1], v.y); hi[2] = fmaxf(hi[2], v.z);
      } else {
        cubql_box3f bx = load_box(c.boxes, __float_as_uint(v.w));
        for (int d = 0; d < 3; d++) { lo[d] = fminf(lo[d], bx.lower[d]); hi[d] = fmaxf(hi[d], bx.upper[d]); }
      }
    }
    for (int d = 0; d < 3; d++) { lo[d] = wave_min_dpp(lo[d]); hi[d] = wave_max_dpp(hi[d]); }
    store_node_wave(loc + rootMe, lo, hi, pack_oc(gb, m));
  } else if (m <= 64) {
    float4 v = lane < m ? gItems[gb + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
    cubql_box3f bx{};
    if (!PTS && lane < m) bx = load_box(c.boxes, __float_as_uint(v.w));
    reg_subtree<PTS>(c, gb, m, v.x, v.y, v.z, __float_as_uint(v.w), bx.lower[0], bx.lower[1],
                     bx.lower[2], bx.upper[0], bx.upper[1], bx.upper[2], rootMe, nLocal, loc);
  } else {
    // more than 64 prims: split in LDS until children fit one wave
    {
      // all loads in flight before the first LDS write
      constexpr int kPer = kSmall / kWave;
      float4 t[kPer];
#pragma unroll
      for (int k = 0; k < kPer; k++) {
        const uint32_t i = (uint32_t)k * kWave + lane;
        if (i < m) t[k] = gItems[gb + i];
      }
#pragma unroll
      for (int k = 0; k < kPer; k++) {
        const uint32_t i = (uint32_t)k * kWave + lane;
        if (i < m) sIt[i] = t[k];
      }
    }
    int sp = 1;
    if (lane == 0) sFr[0] = fr_pack(0u, m, rootMe);
    sync_wave_lds();
    while (sp > 0) {
      --sp;
      const uint32_t f = sFr[sp];
      const uint32_t me = f >> 18;
      const uint32_t b = f & 511u, e = (f >> 9) & 511u;
      const uint32_t n = e - b;
      if (n <= 64) {
        float4 v = lane < n ? sIt[b + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
        cubql_box3f bx{};
        if (!PTS && lane < n) bx = load_box(c.boxes, __float_as_uint(v.w));
        reg_subtree<PTS>(c, gb + b, n, v.x, v.y, v.z, __float_as_uint(v.w), bx.lower[0],
                         bx.lower[1], bx.lower[2], bx.upper[0], bx.upper[1], bx.upper[2], me,
                         nLocal, loc);
        continue;
      }
      // bounds on enc_cmp integers (fused integer DPP reductions)
      int32_t cle[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, che[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
      int32_t ble[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, bhe[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
      for (uint32_t i = b + lane; i < e; i += kWave) {
        const float4 v = sIt[i];
        const int32_t ex = enc_cmp(v.x), ey = enc_cmp(v.y), ez = enc_cmp(v.z);
        cle[0] = min(cle[0], ex); cle[1] = min(cle[1], ey); cle[2] = min(cle[2], ez);
        che[0] = max(che[0], ex); che[1] = max(che[1], ey); che[2] = max(che[2], ez);
        if (!PTS) {
          cubql_box3f bx = load_box(c.boxes, __float_as_uint(v.w));
          for (int d = 0; d < 3; d++) {
            ble[d] = min(ble[d], enc_cmp(bx.lower[d]));
            bhe[d] = max(bhe[d], enc_cmp(bx.upper[d]));
          }
        }
      }
      float clo[3], chi[3], blo[3] = {0.f, 0.f, 0.f}, bhi[3] = {0.f, 0.f, 0.f};
      for (int d = 0; d < 3; d++) { clo[d] = dec_f(wave_min_i32(cle[d])); chi[d] = dec_f(wave_max_i32(che[d])); }
      if (!PTS)
        for (int d = 0; d < 3; d++) { blo[d] = dec_f(wave_min_i32(ble[d])); bhi[d] = dec_f(wave_max_i32(bhe[d])); }
      const float *nlo = PTS ? clo : blo;
      const float *nhi = PTS ? chi : bhi;
      float olo[3] = {nlo[0], nlo[1], nlo[2]}, ohi[3] = {nhi[0], nhi[1], nhi[2]};
      if (n <= T) {
        for (uint32_t i = b + lane; i < e; i += kWave) c.primIDs[gb + i] = __float_as_uint(sIt[i].w);
        store_node_wave(loc + me, olo, ohi, pack_oc(gb + b, n));
        continue;
      }
      int dim;
      float pos;
      choose_split(clo, chi, dim, pos);
      // partition in place: the node's items (<= kSmall/64 per lane) are
      // read into registers first; a wave's LDS accesses run in program
      // order, so the writes below cannot overtake those reads
      constexpr int kPer = kSmall / kWave;
      float4 v[kPer];
      uint32_t dst[kPer];
      bool act[kPer];
      uint32_t nL = 0, nR = 0;
#pragma unroll
      for (int k = 0; k < kPer; k++) {
        const uint32_t i = b + (uint32_t)k * kWave + lane;
        act[k] = i < e;
        v[k] = act[k] ? sIt[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        const bool isL = act[k] && (coord(v[k], dim) < pos);
        const uint64_t mL = __ballot(isL);
        const uint64_t mR = __ballot(act[k] && !isL);
        dst[k] = isL ? b + nL + mbcnt(mL) : e - 1 - (nR + mbcnt(mR));
        nL += (uint32_t)__popcll(mL);
        nR += (uint32_t)__popcll(mR);
      }
      uint32_t mid;
      if (nL != 0 && nR != 0) mid = b + nL;
      else mid = (uint32_t)(((uint64_t)gb + b + gb + e) / 2 - gb);  // spatialMedian.h:107
      sync_wave_lds();
#pragma unroll
      for (int k = 0; k < kPer; k++)
        if (act[k]) sIt[dst[k]] = v[k];
      const uint32_t child = nLocal;
      nLocal += 2;
      store_node_wave(loc + me, olo, ohi, pack_oc(child, 0));
      if (lane == 0) {
        sFr[sp] = fr_pack(mid, e, child + 1);
        sFr[sp + 1] = fr_pack(b, mid, child);
      }
      sp += 2;
      sync_wave_lds();
    }
  }
}

// A pending piece of a medium segment: range [b, e) relative to the
// segment, local id, item buffer, and (when known) its centroid and box
// bounds as enc_cmp integers (lo xyz, hi xyz).
struct MedFrame {
  uint32_t b, e, me, buf;
  int32_t cb[6];
  int32_t bb[6];
};
static_assert(sizeof(MedFrame) == 64, "MedFrame");
constexpr int kMedChunk = 4;  // items per lane per round trip of the medium passes

__device__ __forceinline__ void bounds_init(int32_t (&v)[6]) {
  v[0] = v[1] = v[2] = INT32_MAX;
  v[3] = v[4] = v[5] = INT32_MIN;
}
__device__ __forceinline__ void bounds_add(int32_t (&v)[6], int32_t x, int32_t y, int32_t z, int32_t X,
                                           int32_t Y, int32_t Z) {
  v[0] = min(v[0], x); v[1] = min(v[1], y); v[2] = min(v[2], z);
  v[3] = max(v[3], X); v[4] = max(v[4], Y); v[5] = max(v[5], Z);
}
__device__ __forceinline__ void bounds_wave(int32_t (&v)[6]) {
  for (int d = 0; d < 3; d++) { v[d] = wave_min_i32(v[d]); v[d + 3] = wave_max_i32(v[d + 3]); }
}

// Centroid and box bounds (wave-reduced) of items [lo, hi) of src.
template <bool PTS>
__device__ void med_bounds(const Ctx &c, const float4 *src, uint32_t lo, uint32_t hi, int32_t (&cb)[6],
                           int32_t (&bb)[6]) {
  const uint32_t lane = lane_id();
  bounds_init(cb);
  bounds_init(bb);
  for (uint32_t c0 = lo; c0 < hi; c0 += kMedChunk * kWave) {
    float4 v[kMedChunk];
    bool act[kMedChunk];
#pragma unroll
    for (int k = 0; k < kMedChunk; k++) {
      const uint32_t i = c0 + (uint32_t)k * kWave + lane;
      act[k] = i < hi;
      v[k] = act[k] ? src[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int k = 0; k < kMedChunk; k++) {
      if (!act[k]) continue;
      const int32_t ex = enc_cmp(v[k].x), ey = enc_cmp(v[k].y), ez = enc_cmp(v[k].z);
      bounds_add(cb, ex, ey, ez, ex, ey, ez);
      if (!PTS) {
        const cubql_box3f bx = load_box(c.boxes, __float_as_uint(v[k].w));
        bounds_add(bb, enc_cmp(bx.lower[0]), enc_cmp(bx.lower[1]), enc_cmp(bx.lower[2]),
                   enc_cmp(bx.upper[0]), enc_cmp(bx.upper[1]), enc_cmp(bx.upper[2]));
      }
    }
  }
  bounds_wave(cb);
  if (!PTS) bounds_wave(bb);
}

// One wave finishes a small segment.  Pieces of more than kSmall prims are
// split in global memory: a partition pass from the piece's buffer into the
// other one (left stable, right reversed, one-sided -> middle, as
// everywhere) that also gathers both sides' centroid bounds, so a child piece
// of point input needs no bounds pass of its own unless the split was
// one-sided; pieces that fit
// kSmall are finished at once, larger ones wait on a stack that never holds
// more than kMedium / (kSmall + 1) disjoint pieces.  Local ids are assigned
// in that order.
template <bool PTS>
__device__ void small_segment(const Ctx &c, const SmallSeg &sg, uint32_t segIdx, float4 *sIt, uint32_t *sFr,
                              MedFrame *sMed) {
  const uint32_t lane = lane_id();
  const uint32_t T = c.T;
  const uint32_t gb = sg.begin;
  const uint32_t m = sg.end - sg.begin;
  Node *loc = c.loc + 2 * (size_t)gb;
  uint32_t nLocal = 1;
  auto piece = [&](uint32_t buf, uint32_t b, uint32_t n, uint32_t me) {
    small_sub<PTS>(c, c.items[buf], gb + b, n, me, nLocal, loc, sIt, sFr);
    sync_wave_lds();
  };
  if (m <= (uint32_t)kSmall || m <= T) {
    piece(sg.buf, 0u, m, 0u);
  } else {
    int sp = 1;
    bool known = false;  // bounds of the popped frame
    if (lane == 0) {
      MedFrame f0{};
      f0.b = 0u; f0.e = m; f0.me = 0u; f0.buf = sg.buf | 2u;  // bit 1: bounds unknown
      sMed[0] = f0;
    }
    sync_wave_lds();
    while (sp > 0) {
      --sp;
      const MedFrame f = sMed[sp];
      const uint32_t b = f.b, e = f.e, me = f.me, bf = f.buf & 1u;
      known = (f.buf & 2u) == 0;
      const float4 *src = c.items[bf];
      float4 *dst = c.items[bf ^ 1u];
      int32_t cb[6], bb[6];
      if (known) {
        for (int j = 0; j < 6; j++) { cb[j] = f.cb[j]; bb[j] = f.bb[j]; }
      } else {
        med_bounds<PTS>(c, src, gb + b, gb + e, cb, bb);
      }
      float clo[3], chi[3], olo[3], ohi[3];
      for (int d = 0; d < 3; d++) {
        clo[d] = dec_f(cb[d]); chi[d] = dec_f(cb[d + 3]);
        olo[d] = PTS ? clo[d] : dec_f(bb[d]);
        ohi[d] = PTS ? chi[d] : dec_f(bb[d + 3]);
      }
      int dim;
      float pos;
      choose_split(clo, chi, dim, pos);
      const int32_t posE = enc_cmp(pos);
      // partition pass; per-lane centroid bounds of both sides (point input:
      // the children's boxes; boxes of other input take a bounds pass, which
      // keeps the kernel's register count at the small pieces' occupancy)
      int32_t sc[2][6];
      bounds_init(sc[0]); bounds_init(sc[1]);
      uint32_t nL = 0, nR = 0;
      for (uint32_t c0 = b; c0 < e; c0 += kMedChunk * kWave) {
        float4 v[kMedChunk];
        bool act[kMedChunk];
#pragma unroll
        for (int k = 0; k < kMedChunk; k++) {
          const uint32_t i = c0 + (uint32_t)k * kWave + lane;
          act[k] = i < e;
          v[k] = act[k] ? src[gb + i] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int k = 0; k < kMedChunk; k++) {
          const int32_t ex = enc_cmp(v[k].x), ey = enc_cmp(v[k].y), ez = enc_cmp(v[k].z);
          const int32_t ec = dim == 0 ? ex : (dim == 1 ? ey : ez);
          const bool isL = act[k] && ec < posE;
          const uint64_t mL = __ballot(isL);
          const uint64_t mR = __ballot(act[k] && !isL);
          if (act[k]) {
            const uint32_t d = isL ? b + nL + mbcnt(mL) : e - 1 - (nR + mbcnt(mR));
            dst[gb + d] = v[k];
            const int s = isL ? 0 : 1;
            // (select both sides' updates: no divergent array indexing)
            for (int q = 0; q < 2; q++)
              if (s == q) bounds_add(sc[q], ex, ey, ez, ex, ey, ez);
          }
          nL += (uint32_t)__popcll(mL);
          nR += (uint32_t)__popcll(mR);
        }
      }
      const bool oneSided = nL == 0 || nR == 0;
      const uint32_t mid = !oneSided ? b + nL
                                     : (uint32_t)(((uint64_t)gb + b + gb + e) / 2 - gb);  // spatialMedian.h:107
      const uint32_t child = nLocal;
      nLocal += 2;
      store_node_wave(loc + me, olo, ohi, pack_oc(child, 0));
      // the pieces are read back below, by this wave
      __threadfence_block();
      for (int side = 0; side < 2; side++) {
        const uint32_t cb0 = side ? mid : b, ce0 = side ? e : mid;
        if (ce0 - cb0 <= (uint32_t)kSmall || ce0 - cb0 <= T) {
          piece(bf ^ 1u, cb0, ce0 - cb0, child + side);
        } else {
          MedFrame g{};
          const bool carry = PTS && !oneSided;
          g.b = cb0; g.e = ce0; g.me = child + side; g.buf = (bf ^ 1u) | (carry ? 0u : 2u);
          if (carry) {
            // (wave-uniform branch: every lane joins the reductions)
            int32_t rc[6];
            for (int j = 0; j < 6; j++) rc[j] = side ? sc[1][j] : sc[0][j];
            bounds_wave(rc);
            for (int j = 0; j < 6; j++) g.cb[j] = g.bb[j] = rc[j];
          }
          if (lane == 0) sMed[sp] = g;
          sp++;
          sync_wave_lds();
        }
      }
    }
  }
  if (lane == 0) c.smallCount[segIdx] = nLocal;
}

__global__ __launch_bounds__(kWave) void k_small(Ctx c) {
  __shared__ float4 sIt[kSmall];
  __shared__ uint32_t sFr[kSmall + 2];
  __shared__ MedFrame sMed[kMedium / (kSmall + 1) + 1];
  const bool pts = c.ctr->allPoints != 0;
  const uint32_t nSmall = c.ctr->numSmall;
  for (uint32_t w = blockIdx.x; w < nSmall; w += gridDim.x) {
    const SmallSeg sg = c.small[w];
    if (pts) small_segment<true>(c, sg, w, sIt, sFr, sMed);
    else small_segment<false>(c, sg, w, sIt, sFr, sMed);
    sync_wave_lds();
  }
}

// Exclusive scan of (count-1) over small segments: global id of local node
// j >= 1 of segment s is smallBase[s] + j - 1.
// global ids of the small segments' local nodes (the final node count goes
// to numNodesOut: every workgroup reads nodeCount as its base)
__global__ __launch_bounds__(1024) void k_small_scan(Ctx c, uint32_t tag) {
  __shared__ uint32_t ss[1024 / kWave + 2];
  count_scan_tile(c.smallCount, c.smallBase, c.ctr->numSmall, c.ctr->nodeCount, c.tileStatus, tag,
                  &c.ctr->numNodesOut, ss);
}
__global__ __launch_bounds__(kWave) void k_small_emit(Ctx c, Node *out) {
  const uint32_t nSmall = c.ctr->numSmall;
  for (uint32_t w = blockIdx.x; w < nSmall; w += gridDim.x) {
    const SmallSeg sg = c.small[w];
    const uint32_t cnt = c.smallCount[w];
    const uint32_t base = c.smallBase[w];
    const Node *loc = c.loc + 2 * (size_t)sg.begin;
    for (uint32_t j = lane_id(); j < cnt; j += kWave) {
      Node nd = loc[j];
      if (oc_count(nd.oc) == 0) nd.oc = pack_oc(base + (uint32_t)oc_offset(nd.oc) - 1, 0);
      out[j == 0 ? sg.node : base + j - 1] = nd;
    }
  }
}

__global__ void k_copy_top(const Node *src, Node *dst, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// Bounds of the level-phase inner nodes, deepest level first.
__global__ void k_refit_top(Ctx c, Node *nodes, int level) {
  const uint32_t b = c.levelStart[level], e = c.levelStart[level + 1];
  for (uint32_t i = b + blockIdx.x * blockDim.x + threadIdx.x; i < e; i += gridDim.x * blockDim.x) {
    Node *nd = &nodes[c.topNodes[i]];
    uint64_t off = oc_offset(nd->oc);
    const Node l = nodes[off], r = nodes[off + 1];
    for (int d = 0; d < 3; d++) {
      nd->lo[d] = fminf(l.lo[d], r.lo[d]);
      nd->hi[d] = fmaxf(l.hi[d], r.hi[d]);
    }
  }
}

__global__ void k_init_counters(Counters *k) {
  if (threadIdx.x == 0) {
    for (int d = 0; d < 3; d++) {
      k->rootCb[d] = enc_f(INFINITY);
      k->rootCb[d + 3] = enc_f(-INFINITY);
    }
    k->allPoints = 1;
  }
}

}  // namespace sm

// pinned host mirror for the few completion checks
static_assert(sizeof(sm::Counters) <= 64 * sizeof(uint32_t), "host counter mirror");
static uint32_t *host_words() {
  static thread_local uint32_t *p = nullptr;
  if (!p) {
    if (hipHostMalloc((void **)&p, 64 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) p = nullptr;
  }
  return p;
}

int sm_build(cubql_bvh3f *bvh, const cubql_box3f *boxes, uint32_t n, uint32_t T, hipStream_t s,
             const cubql_mem_resource *mem) {
  using namespace sm;
  Scratch scr(mem, s);
  Ctx c{};
  c.boxes = boxes;
  c.n = n;
  c.T = T;
  const uint32_t numBlocks = div_up(n, kTile);
  c.maxLarge = n / (kMedium + 1) + 2;
  const uint32_t maxTiles = div_up(n, kTile) + c.maxLarge + 1;
  const uint32_t maxLevels = 4096;
  CUBQL_CHECK(scr.alloc(&c.items[0], n));
  CUBQL_CHECK(scr.alloc(&c.items[1], n));
  CUBQL_CHECK(scr.alloc(&c.nodes, 2 * (size_t)n + 2));
  CUBQL_CHECK(scr.alloc(&c.ctr, 1));
  CUBQL_CHECK(scr.alloc(&c.blockCount, numBlocks + 1));
  CUBQL_CHECK(scr.alloc(&c.segs[0], c.maxLarge));
  CUBQL_CHECK(scr.alloc(&c.segs[1], c.maxLarge));
  const uint32_t maxEmitBlocks = div_up(c.maxLarge, kEmitBlock);
  for (int p = 0; p < 2; p++) {
    CUBQL_CHECK(scr.alloc(&c.childCb[p], 12 * (size_t)c.maxLarge * kCbCopies));
    CUBQL_CHECK(scr.alloc(&c.segLeft[p], c.maxLarge));
    CUBQL_CHECK(scr.alloc(&c.tileStart[p], c.maxLarge + 1));
    CUBQL_CHECK(scr.alloc(&c.tileSeg[p], maxTiles));
    CUBQL_CHECK(scr.alloc(&c.segDim[p], c.maxLarge));
    CUBQL_CHECK(scr.alloc(&c.segPos[p], c.maxLarge));
  }
  c.maxEmitBlocks = maxEmitBlocks;
  CUBQL_CHECK(scr.alloc(&c.emitStatus, 4 * (size_t)maxEmitBlocks));
  CUBQL_CHECK(scr.alloc(&c.tickets, kTicketStride * kTicketCounters));
  CUBQL_CHECK(scr.alloc(&c.tileStatus, maxTiles));
  CUBQL_CHECK(scr.alloc(&c.small, n + 1));
  CUBQL_CHECK(scr.alloc(&c.topNodes, n + 1));
  CUBQL_CHECK(scr.alloc(&c.levelStart, maxLevels + 2));
  CUBQL_CHECK(scr.alloc(&c.levelItems, maxLevels + 2));
  CUBQL_CHECK(scr.alloc(&c.blockBounds, 8 * (size_t)numBlocks + 8));
  CUBQL_CHECK(scr.alloc(&c.loc, 2 * (size_t)n + 64));
  CUBQL_CHECK(scr.alloc(&c.smallCount, (size_t)n + 1));
  CUBQL_CHECK(scr.alloc(&c.smallBase, (size_t)n + 1));

  uint32_t *primIDs = nullptr;
  CUBQL_CHECK(mem_alloc(mem, (void **)&primIDs, (size_t)n * sizeof(uint32_t), s));
  c.primIDs = primIDs;
  auto fail = [&](int st) {
    mem_free(mem, primIDs, s);
    return st;
  };

  k_init_counters<<<1, 64, 0, s>>>(c.ctr);
  int st = fill_u32((uint32_t *)c.tileStatus, 0u, 2 * (size_t)maxTiles, s);
  if (!st) st = fill_u32((uint32_t *)c.emitStatus, 0u, 8 * (size_t)maxEmitBlocks, s);
  if (st) return fail(st);
  k_count<<<numBlocks, kBlock, 0, s>>>(c);
  k_count_scan<<<1, 1024, 0, s>>>(c, numBlocks);
  k_write_items<<<numBlocks, kBlock, 0, s>>>(c);
  if (hipGetLastError() != hipSuccess) return fail((set_error("init launch failed"), CUBQL_ERR_HIP));

  uint32_t *hw = host_words();
  if (!hw) return fail((set_error("hipHostMalloc failed"), CUBQL_ERR_HIP));
  int dev = 0, numCU = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&numCU, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    numCU = 256;
  // at least one workgroup per ticket counter, so every counter's tiles run
  // (the partition is persistent: one resident wave of workgroups)
  int partPerCU = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&partPerCU, k_partition, kBlock, 0) != hipSuccess || partPerCU < 1)
    partPerCU = 1;
  const uint32_t partGrid =
      std::max<uint32_t>(kTicketCounters, std::min<uint32_t>(maxTiles, (uint32_t)numCU * (uint32_t)partPerCU));
  int level = 0;
  std::vector<std::pair<int, int>> partSlots;  // (profile slot, level)
  // The first batch covers the depth a well-spread input needs (halving
  // down to kMedium prims, plus one level of slack): running a level past
  // the tree's depth costs two near-empty launches, a host check ~20 us.
  int batch = 1;
  for (uint64_t m = (uint64_t)kMedium; m < n; m *= 2) batch++;
  batch = std::max(kLevelsPerSync, batch);
  while (true) {
    for (int i = 0; i < batch && level < (int)maxLevels; i++, level++) {
      int ps = prof_begin(s);
      k_partition<<<partGrid, kBlock, 0, s>>>(c, level);
      prof_end(ps, s, "sm_partition");
      if (ps >= 0) partSlots.push_back({ps, level});
      k_emit<<<maxEmitBlocks, kEmitBlock, 0, s>>>(c, level);
    }
    if (hipGetLastError() != hipSuccess) return fail((set_error("level launch failed"), CUBQL_ERR_HIP));
    if (hipMemcpyAsync(hw, c.ctr, sizeof(Counters), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return fail((set_error("level sync failed"), CUBQL_ERR_HIP));
    const Counters *hc = (const Counters *)hw;
    // open segments of the next level to run
    if (hc->numLarge[level & 1] == 0) break;
    batch = kLevelsPerSync;
    if (level >= (int)maxLevels) return fail((set_error("tree deeper than %u levels", maxLevels), CUBQL_ERR_INTERNAL));
  }
  const Counters hc0 = *(const Counters *)hw;
  const uint32_t numSmall = hc0.numSmall;
  const uint32_t numValid = hc0.numValid;
  const uint32_t numTop = hc0.numTop;
  const uint32_t topCount = hc0.nodeCount;  // ids [0, topCount) belong to the level phase
  int smallSlot = -1;
  // one workgroup per segment: the hardware dispatcher balances the uneven
  // segment costs far better than a fixed round-robin over resident waves
  const uint32_t smallGrid = std::max<uint32_t>(1u, numSmall);
  if (numSmall) {
    smallSlot = prof_begin(s);
    k_small<<<smallGrid, kWave, 0, s>>>(c);
    prof_end(smallSlot, s, "sm_small");
    // (status tag beyond every level's)
    k_small_scan<<<div_up(numSmall, kCountScanTile), 1024, 0, s>>>(c, maxLevels + 2);
  }
  if (hipGetLastError() != hipSuccess) return fail((set_error("small-phase launch failed"), CUBQL_ERR_HIP));
  if (hipMemcpyAsync(hw, c.ctr, sizeof(Counters), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail((set_error("final sync failed"), CUBQL_ERR_HIP));
  const uint32_t numNodes = numValid ? (numSmall ? ((const Counters *)hw)->numNodesOut : topCount) : 0;
  if (prof_on()) {
    // algorithmic bytes: partition reads and writes one 16-byte item per
    // prim of the level; the small phase reads each item once, writes its
    // primID and the nodes it creates (gathering boxes for non-point input)
    std::vector<uint32_t> li(level + 1, 0);
    if (hipMemcpy(li.data(), c.levelItems, sizeof(uint32_t) * level, hipMemcpyDeviceToHost) == hipSuccess) {
      for (auto &ps : partSlots) prof_set(ps.first, 32.0 * li[ps.second], li[ps.second]);
    }
    const bool pts = hc0.allPoints != 0;
    const double smallNodes = (double)numNodes - (double)topCount + (double)numSmall;
    prof_set(smallSlot, (16.0 + 4.0 + (pts ? 0.0 : 24.0)) * numValid + 32.0 * smallNodes, numValid);
    prof_flush();
  }
  Node *nodes = nullptr;
  st = mem_alloc(mem, (void **)&nodes, std::max<size_t>(numNodes, 2) * sizeof(Node), s);
  if (st) return fail(st);
  if (numNodes) {
    k_copy_top<<<std::max<uint32_t>(1u, std::min<uint32_t>(div_up(topCount, 256), 1024)), 256, 0, s>>>(
        c.nodes, nodes, topCount);
    if (numSmall) k_small_emit<<<smallGrid, kWave, 0, s>>>(c, nodes);
    // (point input: the emit already wrote the level-phase boxes)
    if (numTop && !hc0.allPoints)
      for (int l = level - 1; l >= 0; l--) k_refit_top<<<64, 256, 0, s>>>(c, nodes, l);
    if (hipGetLastError() != hipSuccess) {
      mem_free(mem, nodes, s);
      return fail((set_error("emit launch failed"), CUBQL_ERR_HIP));
    }
  }
  bvh->nodes = (cubql_node3f *)nodes;
  bvh->numNodes = numNodes;
  bvh->primIDs = primIDs;
  bvh->numPrims = n;
  return CUBQL_OK;
}

}  // namespace cubql
