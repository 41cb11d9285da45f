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
//   * one pass splits every open node two tree levels deep: a read-only
//     bounds pass (k_bounds) and a partition (k_pair) that moves each prim
//     straight to its grandchild's range in the CPU order (left stable,
//     right reversed, per level), using decoupled look-back over 2048-item
//     tiles; ranges that come out reversed carry a flag;
//   * nodes of <= kMedium prims are finished by one wavefront each (pieces
//     of <= kSmall prims in LDS), which also computes all their boxes;
//   * the pass loop runs on the device: kernels read their work counts from
//     device memory, the host checks for completion once per batch of
//     passes.
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
// Segments of up to kMedium prims leave the pass loop too: their wave
// splits them in global memory (between the two item buffers) until the
// pieces fit kSmall.
constexpr int kMedium = 1024;
constexpr int kPassesPerSync = 4;
// Tiles are claimed from kTicketCounters interleaved counters (counter c
// hands out tiles c, c+kTicketCounters, ...) and per-segment counts and
// bounds are accumulated into kCbCopies copies: both spread same-address
// atomics, which serialise at one memory-side atomic unit per address.
constexpr uint32_t kTicketCounters = 8;
constexpr uint32_t kTicketStride = 32;  // words between counters
constexpr uint32_t kCbCopies = 8;
// pass A accumulators per segment and copy: six encoded bound sets (left
// side, right side, the halves of a one-sided left split, the halves of a
// one-sided right split) and the left count
constexpr uint32_t kAccA = 40;  // words (36 bounds, count, padding)
// pass B accumulators per segment and copy: the bounds of the four
// grandchild classes
constexpr uint32_t kAccB = 24;
// Pass B gathers the grandchild bounds of a split child only when the child
// has more than kLazyBounds prims: a smaller child's grandchildren are
// nearly always small (their own phase computes their boxes), and the emit
// reduces the bounds of an open one from its prims.
constexpr uint32_t kLazyBounds = 4 * kMedium;
__host__ __device__ __forceinline__ bool gather_bounds(uint32_t split, uint32_t childSize) {
  return split != 0 && childSize > kLazyBounds;
}

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
  uint32_t baseNode[2], baseSmall[2], baseTop[2];  // each pass's bases, by pass parity (emit)
  uint32_t spinFault;  // a look-back wait exceeded its bound (build_common.h)
};

// An open node with > kMedium prims.  [begin, end) is its range in the CPU
// recursion's order; its items sit at [pbeg, pbeg + end - begin) of the
// pass's item buffer, in that order or (rev) in reverse.
struct Seg {
  uint32_t begin, end, node, pbeg;
  int32_t cb[6];      // encoded centroid bounds
  uint32_t rev, pad;
};
static_assert(sizeof(Seg) == 48, "Seg");

// A pass's plan for one segment S, from its pass-A counts: S's split (mode
// 0: by its plane, children [begin, mid) and [mid, end); mode 1 / 2: the
// plane left every prim on the left / right side and the children are the
// halves of the range) and, per child, whether it is split in the same pass
// and by which plane.
struct PairPlan {
  uint32_t begin, end, pbeg, rev;
  uint32_t mid, mode, dim, pos;
  uint32_t split[2], cdim[2];
  uint32_t cpos[2], pad[2];
  int32_t ccb[2][6];  // children's encoded centroid bounds
  uint32_t pad2[4];
};
static_assert(sizeof(PairPlan) == 128, "PairPlan");

// Order of child `side` of a plan relative to S's order: the right child of
// a split by the plane holds S's right-side prims in reverse (the CPU order),
// the halves of a one-sided split keep the order of the partitioned range.
__host__ __device__ __forceinline__ uint32_t child_rev(uint32_t mode, uint32_t side) {
  return mode == 0 ? side : (mode == 2 ? 1u : 0u);
}

struct SmallSeg {     // a node finished by one wave
  uint32_t begin, end, node, buf;  // range (CPU order), node, item buffer
  uint32_t pbeg, rev, pad0, pad1;  // first slot of its items, reverse storage
};
static_assert(sizeof(SmallSeg) == 32, "SmallSeg");

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
  // per-pass arrays, double-buffered by pass parity: the emit of pass p
  // reads the [p&1] copies and prepares the [(p+1)&1] ones
  Seg *segs[2];            // open segments of a pass
  uint32_t *tileStart[2];  // [maxLarge+1] first tile of each segment
  uint32_t *tileSeg[2];    // [maxTiles] segment of each tile
  int32_t *segDim[2];      // [maxLarge] split plane
  float *segPos[2];
  int32_t *accA;           // [maxLarge][kCbCopies][kAccA] pass A sums (k_bounds)
  int32_t *accB;           // [maxLarge][kCbCopies][kAccB] pass B bounds (k_pair)
  PairPlan *plan;          // [maxLarge] (k_plan)
  uint32_t *classTot;      // [maxLarge][4] grandchild class sizes (k_pair)
  uint32_t *tileLeft;      // [maxTiles] left-side prims of each tile (k_bounds)
  uint32_t *tilePre;       // [maxTiles] left-side prims before each tile, over all tiles (k_tile_scan)
  uint64_t *emitStatus;    // [5][maxEmitBlocks] emit look-back status per count
  uint32_t maxEmitBlocks;
  uint64_t *tileStatus;    // [4][maxTiles] look-back status per class
  uint32_t maxTiles;
  SmallSeg *small;         // [numValid]
  uint32_t *topNodes;      // [numValid] inner nodes of the pass phase
  uint32_t *levelStart;    // [2*maxPasses+1] start of each node list in topNodes
  uint32_t *levelItems;    // [maxPasses] prims moved by each pass
  uint32_t maxLarge;
};

// Accumulator copies in use for a segment of n prims: enough to keep the
// same-address atomics of its tiles' workgroups (one flush per workgroup
// and segment) in short queues, few enough that its readers merge little.
__host__ __device__ __forceinline__ uint32_t acc_copies(uint32_t n) {
  const uint32_t tiles = div_up(n, (uint32_t)kTile);
  return tiles >= 64 * kCbCopies ? kCbCopies : (tiles < 128 ? 1u : tiles / 64);
}

// Accumulator identities: min slots +inf, max slots -inf (bound sets of six
// words, lo xyz then hi xyz); pass A's count word (36) and padding are 0.
__device__ __forceinline__ int32_t acc_identity(uint32_t w) {
  return (w % 6) < 3 ? enc_f(INFINITY) : enc_f(-INFINITY);
}
static_assert(kAccA == 40 && kAccB == 24, "accumulator layout");
// 16-byte word group q of a segment's pass-A accumulators (kCbCopies * 10
// groups, copies contiguous), cleared
__device__ __forceinline__ int4 acc_a_clear(uint32_t q) {
  const uint32_t w = 4 * (q % 10);
  if (w >= 36) return make_int4(0, 0, 0, 0);
  return make_int4(acc_identity(w), acc_identity(w + 1), acc_identity(w + 2), acc_identity(w + 3));
}
__device__ __forceinline__ int4 acc_b_clear(uint32_t q) {
  const uint32_t w = 4 * (q % 6);
  return make_int4(acc_identity(w), acc_identity(w + 1), acc_identity(w + 2), acc_identity(w + 3));
}


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
      s.begin = 0; s.end = run; s.node = 0; s.pbeg = 0; s.rev = 0; s.pad = 0;
      for (int j = 0; j < 6; j++) s.cb[j] = k->rootCb[j];
      c.segs[0][0] = s;
      k->numLarge[0] = 1;
      // the root's plane
      float lo[3], hi[3];
      for (int d = 0; d < 3; d++) { lo[d] = dec_f(s.cb[d]); hi[d] = dec_f(s.cb[d + 3]); }
      int dim;
      float pos;
      choose_split(lo, hi, dim, pos);
      c.segDim[0][0] = dim;
      c.segPos[0][0] = pos;
      const uint32_t tiles = div_up(run, kTile);
      c.tileStart[0][0] = 0;
      c.tileStart[0][1] = tiles;
      k->numTiles = tiles;
      c.levelItems[0] = run;
    } else if (run > 0) {
      SmallSeg s{0, run, 0, 0, 0, 0, 0, 0};
      c.small[0] = s;
      k->numSmall = 1;
    }
    // pass 0's bases for the emit
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

// ------------------------------------------------------------- pair passes
// Every pass splits each open segment S two tree levels deep, with the
// result of the CPU recursion (cpu/spatialMedian.h:81-114), which
// partitions a node's range in place -- left-side prims first in their
// order, right-side prims after them in reverse order -- and splits a node
// whose plane leaves one side empty at the middle position:
//   * k_bounds (pass A) reads S's items once and accumulates the left count
//     and the centroid bounds of both sides of S's plane (and of the position
//     halves a one-sided split would take);
//   * k_plan turns them into S's children: ranges, bounds and planes;
//   * k_pair (pass B) reads S's items again and writes each straight to the
//     range of its grandchild;
//   * k_emit2 writes the nodes, the next pass's segments and tile map, and
//     the small-segment list.
// The items cross HBM three times per two levels instead of four.  A
// grandchild range can only be filled from the two ends of its parent's
// range (where they meet is known when the pass ends), so the children of a
// node whose order is the reverse of S's -- the right child -- come out
// reversed: a segment carries that as a flag and readers walk it backwards.


// Items [j0, j0 + cnt) of a segment's order (striped: item jj of thread t is
// j0 + jj*kBlock + t); lanes past the end re-read the last item.  (Slots
// are clamped to the buffer, lim items.)
__device__ __forceinline__ void load_tile(const float4 *src, uint32_t pbeg, uint32_t n, uint32_t rev,
                                          uint32_t j0, uint32_t cnt, uint32_t lim, float4 (&it)[kItems]) {
#pragma unroll
  for (int jj = 0; jj < kItems; jj++) {
    const uint32_t j = j0 + min((uint32_t)(jj * kBlock) + threadIdx.x, cnt - 1);
    it[jj] = src[min(rev ? pbeg + n - 1 - j : pbeg + j, lim - 1)];
  }
}

// Pass A: each workgroup walks a contiguous run of the pass's tiles (no tile
// waits on another).  It accumulates per thread while its tiles belong to
// one segment and adds its sums to the segment's accumulators (one of
// kCbCopies copies) when the segment changes, so a segment's slots take
// about one atomic per workgroup and word, not one per tile.
// The next tile's loads are in flight during that flush.  One thread of
// wave 0 reads the tile descriptors (into LDS) and wave 3 issues the
// atomics: a wave's memory counter retires in order, so a wave that waited
// on its descriptor loads behind its own atomics would wait for them too.
struct TileA {
  uint32_t s, n, j0, cnt, pbeg, rev;
  int dim;
  float pos;
};

// bounds on enc_f integers (the accumulators' order): integer min/max, and
// fused DPP steps for the wave reductions
#define CUBQL_ADD6(A, O, X, Y, Z)                                               \
  {                                                                             \
    A[O] = min(A[O], X); A[O + 1] = min(A[O + 1], Y); A[O + 2] = min(A[O + 2], Z); \
    A[O + 3] = max(A[O + 3], X); A[O + 4] = max(A[O + 4], Y); A[O + 5] = max(A[O + 5], Z); \
  }

__global__ __launch_bounds__(kBlock) void k_bounds(Ctx c, int pass) {
  constexpr int NW = kBlock / kWave;
  constexpr uint32_t kAtomicWave = NW - 1;
  __shared__ int32_t sb[36 * NW];
  __shared__ uint32_t sc[NW];
  __shared__ TileA sD[2];
  __shared__ uint32_t sTL[2][NW];  // per-wave left counts of a tile, by tile parity
  const int par = pass & 1;
  const uint32_t numTiles = c.ctr->numTiles;
  const float4 *src = c.items[par];
  const uint32_t lane = lane_id(), wid = threadIdx.x / kWave;
  auto U = [](uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); };
  auto readD = [&](int sl) {
    TileA d;
    d.s = U(sD[sl].s); d.n = U(sD[sl].n); d.j0 = U(sD[sl].j0); d.cnt = U(sD[sl].cnt);
    d.pbeg = U(sD[sl].pbeg); d.rev = U(sD[sl].rev); d.dim = (int)U((uint32_t)sD[sl].dim);
    d.pos = __uint_as_float(U(__float_as_uint(sD[sl].pos)));
    return d;
  };
  auto desc = [&](uint32_t t) {
    TileA d;
    d.s = c.tileSeg[par][t];
    const Seg g = c.segs[par][d.s];
    d.n = g.end - g.begin;
    d.j0 = (t - c.tileStart[par][d.s]) * kTile;
    d.cnt = min(d.n - d.j0, (uint32_t)kTile);
    d.pbeg = g.pbeg;
    d.rev = g.rev;
    d.dim = c.segDim[par][d.s];
    d.pos = c.segPos[par][d.s];
    return d;
  };
  const uint32_t per = div_up(numTiles, gridDim.x);
  uint32_t t = blockIdx.x * per;
  const uint32_t tEnd = min(t + per, numTiles);
  if (t >= tEnd) return;
  int sl = 0;
  if (threadIdx.x == 0) sD[0] = desc(t);
  sync_block_lds();
  TileA d = readD(0);
  float4 it[kItems];
  load_tile(src, d.pbeg, d.n, d.rev, d.j0, d.cnt, c.n, it);
  // sets: left / right of the plane (0, 6); positions < h / >= h, the
  // halves of an all-left split (12, 18); positions >= n-h / < n-h, the
  // halves of an all-right split (24, 30); h = n/2
  int32_t a[36];
  for (int i = 0; i < 36; i++) a[i] = acc_identity(i);
  uint32_t nl = 0;
  bool anyR = false;  // a right-side prim seen since the last flush
  while (true) {
    const uint32_t h = d.n / 2;
    uint32_t wl = 0;  // the wave's left-side prims in this tile
#pragma unroll
    for (int jj = 0; jj < kItems; jj++) {
      const uint32_t q = (uint32_t)(jj * kBlock) + threadIdx.x;
      const float4 p = it[jj];
      const bool lf = q < d.cnt && coord(p, d.dim) < d.pos;
      wl += (uint32_t)__popcll(__ballot(lf));
      if (q < d.cnt) {
        const uint32_t j = d.j0 + q;
        const int32_t ex = enc_f(p.x), ey = enc_f(p.y), ez = enc_f(p.z);
        if (lf) {
          nl++;
          CUBQL_ADD6(a, 0, ex, ey, ez);
        } else {
          anyR = true;
          CUBQL_ADD6(a, 6, ex, ey, ez);
        }
        if (j < h) CUBQL_ADD6(a, 12, ex, ey, ez) else CUBQL_ADD6(a, 18, ex, ey, ez);
        if (j >= d.n - h) CUBQL_ADD6(a, 24, ex, ey, ez) else CUBQL_ADD6(a, 30, ex, ey, ez);
      }
    }
    if (lane == 0) sTL[t & 1][wid] = wl;
    const uint32_t t2 = t + 1;
    const bool more = t2 < tEnd;
    TileA d2 = d;
    if (more) {
      // (the other slot's last reader passed the barrier before this
      // tile's loads)
      if (threadIdx.x == 0) sD[sl ^ 1] = desc(t2);
      sync_block_lds();
      sl ^= 1;
      d2 = readD(sl);
      load_tile(src, d2.pbeg, d2.n, d2.rev, d2.j0, d2.cnt, c.n, it);
    }
    if (!more || d2.s != d.s) {
      // a wave that saw prims on both sides of the plane has shown that the
      // split is not one-sided: the position halves are not needed from it
      const bool both = __ballot(nl != 0) != 0 && __ballot(anyR) != 0;
#pragma unroll
      for (int i = 0; i < 12; i++) a[i] = (i % 6) < 3 ? wave_min_i32(a[i]) : wave_max_i32(a[i]);
      if (both) {
#pragma unroll
        for (int i = 12; i < 36; i++) a[i] = acc_identity(i);
      } else {
#pragma unroll
        for (int i = 12; i < 36; i++) a[i] = (i % 6) < 3 ? wave_min_i32(a[i]) : wave_max_i32(a[i]);
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) nl += __shfl_xor(nl, o, kWave);
      if (lane == 0) {
#pragma unroll
        for (int i = 0; i < 36; i++) sb[i * NW + wid] = a[i];
        sc[wid] = nl;
      }
      sync_block_lds();
      int32_t *acc = c.accA + ((size_t)d.s * kCbCopies + blockIdx.x % acc_copies(d.n)) * kAccA;
      const uint32_t i = threadIdx.x - kAtomicWave * kWave;
      if (wid == kAtomicWave && i < 36) {
        int32_t r = sb[i * NW];
        for (int w = 1; w < NW; w++) r = (i % 6) < 3 ? min(r, sb[i * NW + w]) : max(r, sb[i * NW + w]);
        if (r != acc_identity(i)) {
          if ((i % 6) < 3) atomicMin(&acc[i], r);
          else atomicMax(&acc[i], r);
        }
      } else if (wid == kAtomicWave && i == 36) {
        uint32_t tot = 0;
        for (int w = 0; w < NW; w++) tot += sc[w];
        if (tot) atomicAdd(&acc[36], (int32_t)tot);
      }
      sync_block_lds();
      for (int i = 0; i < 36; i++) a[i] = acc_identity(i);
      nl = 0;
      anyR = false;
    }
    // (a barrier has passed since the counts were written: the next tile's
    // or the flush's; the slot is rewritten two tiles later, past the next
    // tile's barrier)
    if (threadIdx.x == 0) {
      uint32_t tl = 0;
      for (int w = 0; w < NW; w++) tl += sTL[t & 1][w];
      c.tileLeft[t] = tl;
    }
    if (!more) break;
    t = t2;
    d = d2;
  }
}

// Plans of the pass's segments, one wave each, from the pass-A sums; the
// wave also clears the segment's pass-B accumulators.  The first scanBlocks
// workgroups instead scan the tiles' left-side counts over all the pass's
// tiles (pass B takes a tile's offset within its segment as a difference;
// status region 2 of the tile look-back array).
constexpr uint32_t kTileScanPart = 256 * kCountScanPer;
__global__ __launch_bounds__(256) void k_plan(Ctx c, int pass, uint32_t scanBlocks) {
  if (blockIdx.x < scanBlocks) {
    __shared__ uint32_t ss[256 / kWave + 2];
    count_scan_part<256>(c.tileLeft, c.tilePre, c.ctr->numTiles, 0u, c.tileStatus + 2 * (size_t)c.maxTiles,
                         (uint32_t)pass + 1, blockIdx.x, false, nullptr, ss, 0u, &c.ctr->spinFault);
    return;
  }
  const int par = pass & 1;
  const uint32_t n = c.ctr->numLarge[par];
  const uint32_t s = (blockIdx.x - scanBlocks) * (256 / kWave) + threadIdx.x / kWave;
  if (s >= n) return;  // (whole waves)
  const uint32_t lane = lane_id();
  const int32_t *acc = c.accA + (size_t)s * kCbCopies * kAccA;
  const Seg g = c.segs[par][s];
  const uint32_t size = g.end - g.begin;
  const uint32_t cps = acc_copies(size);
  uint32_t nl = lane < cps ? (uint32_t)acc[lane * kAccA + 36] : 0u;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) nl += __shfl_xor(nl, o, kWave);
  const uint32_t mode = nl == size ? 1u : (nl == 0 ? 2u : 0u);
  // lane j < 12: word j of the children's bound sets, over the copies
  int32_t v = acc_identity(lane % 12);
  if (lane < 12)
    for (uint32_t cp = 0; cp < cps; cp++) {
      const int32_t x = acc[cp * kAccA + 12 * mode + lane];
      v = (lane % 6) < 3 ? min(v, x) : max(v, x);
    }
  int32_t cb[12];
#pragma unroll
  for (int j = 0; j < 12; j++) cb[j] = __shfl(v, j, kWave);
  if (lane == 0) {
    PairPlan P{};
    P.begin = g.begin; P.end = g.end; P.pbeg = g.pbeg; P.rev = g.rev;
    P.mid = mode ? g.begin + size / 2 : g.begin + nl;  // spatialMedian.h:107 for one-sided splits
    P.mode = mode;
    P.dim = (uint32_t)c.segDim[par][s];
    P.pos = __float_as_uint(c.segPos[par][s]);
    for (int side = 0; side < 2; side++) {
      const uint32_t sz = side ? g.end - P.mid : P.mid - g.begin;
      P.split[side] = sz > (uint32_t)kMedium && sz > c.T;
      float lo[3], hi[3];
      for (int d = 0; d < 3; d++) {
        P.ccb[side][d] = cb[6 * side + d];
        P.ccb[side][d + 3] = cb[6 * side + d + 3];
        lo[d] = dec_f(cb[6 * side + d]);
        hi[d] = dec_f(cb[6 * side + d + 3]);
      }
      if (P.split[side]) {
        int dim;
        float pos;
        choose_split(lo, hi, dim, pos);
        P.cdim[side] = (uint32_t)dim;
        P.cpos[side] = __float_as_uint(pos);
      }
    }
    c.plan[s] = P;
  }
  if (lane < kCbCopies * kAccB / 4)
    reinterpret_cast<int4 *>(c.accB + (size_t)s * kCbCopies * kAccB)[lane] = acc_b_clear(lane);
  // clean the pass-A slots for the next pass that uses segment index s
  // (every slot starts clean: k_init_acc)
  int4 *a4 = reinterpret_cast<int4 *>(c.accA + (size_t)s * kCbCopies * kAccA);
  for (uint32_t q = lane; q < kCbCopies * kAccA / 4; q += kWave) a4[q] = acc_a_clear(q);
}

// Clears every segment slot's pass-A accumulators once per build.
// One launch before the build: the pass-A accumulators, the look-back
// status words of the passes and emits (nStatus words from tileStatus,
// nEmit from emitStatus) and the counters.
__global__ void k_init_acc(Ctx c, size_t nStatus, size_t nEmit) {
  constexpr uint32_t kGroups = kCbCopies * kAccA / 4;
  int4 *a4 = reinterpret_cast<int4 *>(c.accA);
  const size_t total = (size_t)c.maxLarge * kGroups;
  const size_t g0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x, gs = (size_t)gridDim.x * blockDim.x;
  for (size_t i = g0; i < total; i += gs) a4[i] = acc_a_clear((uint32_t)(i % kGroups));
  uint32_t *ts = reinterpret_cast<uint32_t *>(c.tileStatus);
  for (size_t i = g0; i < nStatus; i += gs) ts[i] = 0u;
  uint32_t *es = reinterpret_cast<uint32_t *>(c.emitStatus);
  for (size_t i = g0; i < nEmit; i += gs) es[i] = 0u;
  if (g0 == 0) {
    Counters *k = c.ctr;
    for (int d = 0; d < 3; d++) {
      k->rootCb[d] = enc_f(INFINITY);
      k->rootCb[d + 3] = enc_f(-INFINITY);
    }
    k->allPoints = 1;
    k->spinFault = 0;
  }
}

// Pass B, software-pipelined over a workgroup's tiles like a one-level
// partition: a workgroup claims tiles in ticket order (so a tile only ever
// waits on tiles claimed by running workgroups), classifies its items into
// the four grandchild classes (child 0 or 1 of S, then that child's left or
// right side; an unsplit child is one class), stages them in LDS in output
// order, resolves the four prefixes by decoupled look-back among the tiles
// of its segment (one wave per class) and copies the staged items out in
// four contiguous runs.  Class (child, left) fills the child's range from
// its start, class (child, right) from its end.  After its look-back the
// workgroup claims its next tile, which is loaded, classified, counted and
// published before this tile's copy-out.  A workgroup holds at most two
// tiles, the later one claimed after the earlier; a tile publishes its class
// counts before it waits on anything but its own loads, so every wait ends
// and the grid drains.
__global__ __launch_bounds__(kBlock, 4) void k_pair(Ctx c, int pass) {
  constexpr int NW = kBlock / kWave;
  static_assert(NW == 4, "one look-back wave per class");
  __shared__ float4 sStage[kTile];
  __shared__ uint32_t sr[4 * kItems * NW + 4];
  __shared__ int32_t sb[24 * NW];
  __shared__ PairPlan sPlan[2];
  __shared__ uint32_t sTile[2][5];  // tile, segment, first tile, end tile, left prims before it
  __shared__ uint32_t sPrefix[4];
  const int par = pass & 1;
  const uint32_t numTiles = c.ctr->numTiles;
  if (numTiles == 0) return;  // a pass past the tree's depth
  const float4 *src = c.items[par];
  float4 *dst = c.items[par ^ 1];
  const uint32_t tag = (uint32_t)pass + 1;
  const uint32_t lane = lane_id(), wid = threadIdx.x / kWave;
  const uint32_t cc = blockIdx.x % kTicketCounters;
  auto fetch = [&](int sl) {  // one thread
    const uint32_t t = cc + kTicketCounters * atomicAdd(&c.tickets[kTicketStride * cc], 1u);
    sTile[sl][0] = t;
    if (t < numTiles) {
      const uint32_t sg = c.tileSeg[par][t];
      sTile[sl][1] = sg;
      sTile[sl][2] = c.tileStart[par][sg];
      sTile[sl][3] = c.tileStart[par][sg + 1];
      sTile[sl][4] = c.tilePre[t] - c.tilePre[sTile[sl][2]];
      sPlan[sl] = c.plan[sg];
    }
  };
  auto U = [](uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); };
  int slot = 0;
  if (threadIdx.x == 0) fetch(0);
  sync_block_lds();
  uint32_t t = U(sTile[0][0]);
  if (t >= numTiles) return;
  float4 it[kItems];
  // class bounds, per thread, over the workgroup's consecutive tiles of one
  // segment; added to the segment's accumulators when the segment changes,
  // by threads 0-23, after the next tile's look-back (a wave's memory
  // counter retires in order: atomics issued just before a look-back would
  // hold up its status loads)
  int32_t cacc[24];
  for (int i = 0; i < 24; i++) cacc[i] = acc_identity(i);
  int32_t pend = 0;
  uint32_t pendSeg = 0;
  bool pendOn = false;
  uint32_t pendCps = 1;
  auto addB = [&](uint32_t sg, uint32_t cps, int32_t r) {
    const uint32_t i = threadIdx.x;
    if (i < 24 && r != acc_identity(i)) {
      int32_t *acc = c.accB + ((size_t)sg * kCbCopies + blockIdx.x % cps) * kAccB;
      if ((i % 6) < 3) atomicMin(&acc[i], r);
      else atomicMax(&acc[i], r);
    }
  };
  // A tile is classified (class 2 * child + side of each item, 4 = none;
  // three bits an item), counted (per-wave class counts and offsets in sr,
  // class totals) and its look-back counts published as soon as its items
  // are in; it is ranked and staged later.  A workgroup counts its next tile
  // before it copies out the current one, so the tiles that look back over
  // the next one do not wait for that copy-out.
  auto classify = [&](int sl, uint32_t tt) {
    const PairPlan &P = sPlan[sl];
    const uint32_t n = U(P.end) - U(P.begin), h = n / 2, mode = U(P.mode);
    const int dim = (int)U(P.dim);
    const float pos = __uint_as_float(U(P.pos));
    const uint32_t split0 = U(P.split[0]), split1 = U(P.split[1]);
    const int cdim0 = (int)U(P.cdim[0]), cdim1 = (int)U(P.cdim[1]);
    const float cpos0 = __uint_as_float(U(P.cpos[0])), cpos1 = __uint_as_float(U(P.cpos[1]));
    const uint32_t j0 = (tt - U(sTile[sl][2])) * kTile;
    const uint32_t cnt = min(n - j0, (uint32_t)kTile);
    uint32_t pk = 0;
#pragma unroll
    for (int jj = 0; jj < kItems; jj++) {
      const uint32_t q = (uint32_t)(jj * kBlock) + threadIdx.x;
      const uint32_t j = j0 + q;
      const uint32_t c1 = mode == 0 ? (coord(it[jj], dim) < pos ? 0u : 1u)
                                    : (mode == 1 ? (j >= h ? 1u : 0u) : (j < n - h ? 1u : 0u));
      const bool sp = c1 ? split1 != 0 : split0 != 0;
      const int cd = c1 ? cdim1 : cdim0;
      const float cpv = c1 ? cpos1 : cpos0;
      const uint32_t c2 = (sp && !(coord(it[jj], cd) < cpv)) ? 1u : 0u;
      pk |= (q < cnt ? 2 * c1 + c2 : 4u) << (3 * jj);
    }
    return pk;
  };
  constexpr int NC = kItems * NW;  // counters per class
  static_assert(4 * NC == 2 * kWave, "two counters per lane");
  auto count4 = [&](uint32_t pk, uint32_t (&tot)[4]) {
#pragma unroll
    for (int jj = 0; jj < kItems; jj++) {
      const uint32_t cl = (pk >> (3 * jj)) & 7u;
#pragma unroll
      for (uint32_t q = 0; q < 4; q++) {
        const uint64_t m = __ballot(cl == q);
        if (lane == 0) sr[(q * kItems + jj) * NW + wid] = (uint32_t)__popcll(m);
      }
    }
    sync_block_lds();
    if (wid == 0) {
      // lane l scans counters 2l, 2l+1; class q owns lanes [16q, 16q+16)
      const uint32_t a = sr[2 * lane], b = sr[2 * lane + 1];
      const uint32_t incl = wave_incl_scan(a + b);
      const uint32_t ex = incl - a - b;
      const uint32_t base = (uint32_t)__shfl((int)ex, (int)(lane & ~15u), kWave);
      sr[2 * lane] = ex - base;
      sr[2 * lane + 1] = ex - base + a;
      if ((lane & 15u) == 15u) sr[4 * NC + (lane >> 4)] = incl - base;
    }
    sync_block_lds();
#pragma unroll
    for (int q = 0; q < 4; q++) tot[q] = U(sr[4 * NC + q]);
  };
  auto publish = [&](int sl, uint32_t tt, const uint32_t (&tot)[4]) {
    const PairPlan &P = sPlan[sl];
    if (wid < 2 && U(wid == 0 ? P.split[0] : P.split[1]) != 0)
      lookback_publish(c.tileStatus + (size_t)wid * c.maxTiles, tt, tt - U(sTile[sl][2]), tag,
                       wid == 0 ? tot[0] : tot[2]);
  };
  uint32_t pk, tot[4];
  {
    const uint32_t n = U(sPlan[0].end) - U(sPlan[0].begin);
    const uint32_t j0 = (t - U(sTile[0][2])) * kTile;
    load_tile(src, U(sPlan[0].pbeg), n, U(sPlan[0].rev), j0, min(n - j0, (uint32_t)kTile), c.n, it);
    pk = classify(0, t);
    count4(pk, tot);
    publish(0, t, tot);
  }
  while (true) {
    const PairPlan &P = sPlan[slot];
    const uint32_t seg = U(sTile[slot][1]), tileFirst = U(sTile[slot][2]), tileEnd = U(sTile[slot][3]);
    const uint32_t begin = U(P.begin), end = U(P.end), mid = U(P.mid), mode = U(P.mode);
    const uint32_t split0 = U(P.split[0]), split1 = U(P.split[1]);
    const uint32_t n = end - begin, h = n / 2;
    const uint32_t k = t - tileFirst;
    const uint32_t j0 = k * kTile;
    const uint32_t ctot[4] = {tot[0], tot[1], tot[2], tot[3]};
    const uint32_t off1 = ctot[0], off2 = off1 + ctot[1], off3 = off2 + ctot[2];
    // (the previous tile's copy-out has read sStage)
    sync_block_lds();
    // rank (same-class items before each item: the wave's ballot prefix
    // plus the wave's offset) and stage in output order
#pragma unroll
    for (int jj = 0; jj < kItems; jj++) {
      const uint32_t cl = (pk >> (3 * jj)) & 7u;
      uint32_t r = 0;
#pragma unroll
      for (uint32_t q = 0; q < 4; q++) {
        const uint64_t m = __ballot(cl == q);
        if (cl == q) r = mbcnt(m);
      }
      if (cl < 4) {
        r += sr[(cl * kItems + jj) * NW + wid];
        const uint32_t idx = cl == 0 ? r : (cl == 1 ? off2 - 1 - r : (cl == 2 ? off2 + r : off3 + ctot[3] - 1 - r));
        sStage[idx] = it[jj];
      }
    }
    // Offsets: the prims of child 0 before this tile are known (pass A's
    // counts, or positions for a one-sided split); only the left classes of
    // split children need a look-back (waves 0 and 1), the right classes
    // follow by difference, an unsplit child is one class.
    {
      const uint32_t q = 2 * wid;  // class (wid, left)
      const bool look = wid < 2 && (wid == 0 ? split0 : split1) != 0;
      uint32_t pre = 0;
      if (look) {
        const uint32_t mine = wid == 0 ? ctot[0] : ctot[2];
        pre = lookback_exclusive_wave(c.tileStatus + (size_t)wid * c.maxTiles, t, k, tag, mine, true,
                                      &c.ctr->spinFault);
        if (lane == 0 && t == tileEnd - 1) c.classTot[4 * (size_t)seg + q] = pre + mine;
      }
      if (lane == 0) {
        if (wid < 2) sPrefix[wid] = pre;
        if (wid == 0) fetch(slot ^ 1);
      }
    }
    sync_block_lds();
    const uint32_t c0pre = mode == 0 ? U(sTile[slot][4]) : (mode == 1 ? min(j0, h) : (j0 > n - h ? j0 - (n - h) : 0u));
    const uint32_t c1pre = j0 - c0pre;
    const uint32_t p0 = split0 ? sPrefix[0] : c0pre;
    const uint32_t p1 = split0 ? c0pre - p0 : 0u;
    const uint32_t p2 = split1 ? sPrefix[1] : c1pre;
    const uint32_t p3 = split1 ? c1pre - p2 : 0u;
    const int nslot = slot ^ 1;
    const uint32_t nt = U(sTile[nslot][0]);
    // the next tile: loaded, classified, counted and published before this
    // tile's copy-out (sr is free: this tile's ranks are done)
    if (nt < numTiles) {
      const PairPlan &Q = sPlan[nslot];
      const uint32_t n2 = U(Q.end) - U(Q.begin);
      const uint32_t j02 = (nt - U(sTile[nslot][2])) * kTile;
      load_tile(src, U(Q.pbeg), n2, U(Q.rev), j02, min(n2 - j02, (uint32_t)kTile), c.n, it);
      pk = classify(nslot, nt);
      count4(pk, tot);
      publish(nslot, nt, tot);
    }
    if (pendOn) {
      addB(pendSeg, pendCps, pend);
      pendOn = false;
    }
    // copy out: staged class q -> its run; bounds of each class on the way
    const uint32_t runOff[4] = {0u, off1, off2, off3};
    const uint32_t runBase[4] = {begin + p0, mid - p1 - ctot[1], mid + p2, end - p3 - ctot[3]};
    const bool gb0 = gather_bounds(split0, mid - begin), gb1 = gather_bounds(split1, end - mid);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint32_t o = runOff[q], cntq = ctot[q], rb = runBase[q];
      if (q < 2 ? gb0 : gb1) {
#pragma unroll 4
        for (uint32_t i = threadIdx.x; i < cntq; i += kBlock) {
          const float4 x = sStage[o + i];
          dst[min(rb + i, c.n - 1)] = x;  // (in range by construction; clamped to the buffer)
          const int32_t ex = enc_f(x.x), ey = enc_f(x.y), ez = enc_f(x.z);
          CUBQL_ADD6(cacc, 6 * q, ex, ey, ez);
        }
      } else {
#pragma unroll 4
        for (uint32_t i = threadIdx.x; i < cntq; i += kBlock) dst[min(rb + i, c.n - 1)] = sStage[o + i];
      }
    }
    if ((gb0 || gb1) && (nt >= numTiles || U(sTile[nslot][1]) != seg)) {
#pragma unroll
      for (int i = 0; i < 24; i++) cacc[i] = (i % 6) < 3 ? wave_min_i32(cacc[i]) : wave_max_i32(cacc[i]);
      if (lane == 0) {
#pragma unroll
        for (int i = 0; i < 24; i++) sb[i * NW + wid] = cacc[i];
      }
      sync_block_lds();
      int32_t r = 0;
      if (threadIdx.x < 24) {
        const uint32_t i = threadIdx.x;
        r = sb[i * NW];
        for (int w = 1; w < NW; w++) r = (i % 6) < 3 ? min(r, sb[i * NW + w]) : max(r, sb[i * NW + w]);
      }
      sync_block_lds();
      if (nt >= numTiles) {
        addB(seg, acc_copies(n), r);
      } else {
        pend = r;
        pendSeg = seg;
        pendCps = acc_copies(n);
        pendOn = true;
      }
      for (int i = 0; i < 24; i++) cacc[i] = acc_identity(i);
    }
    if (nt >= numTiles) break;
    t = nt;
    slot = nslot;
  }
}

// A grandchild of S under child range [cb, ce) (child order rv relative to
// S), side gs: its range in the CPU order, its storage direction and whether
// it stays open.  nCL = the child's left-side count.  A one-sided child
// split takes the halves of the child's range; a reversed child is first
// turned around in place by the emit (otherwise each half would sit in the
// other half's range), so its halves are stored in order.
struct GrandKid {
  uint32_t b, e, rev;
  bool large, oneSided;
};
__device__ __forceinline__ GrandKid grand_kid(uint32_t cb, uint32_t ce, uint32_t nCL, uint32_t rv, uint32_t gs,
                                              uint32_t T) {
  GrandKid g;
  const uint32_t sz = ce - cb;
  g.oneSided = nCL == 0 || nCL == sz;
  const uint32_t mid = g.oneSided ? cb + sz / 2 : cb + nCL;  // spatialMedian.h:107
  g.b = gs ? mid : cb;
  g.e = gs ? ce : mid;
  g.rev = g.oneSided ? 0u : rv;
  g.large = g.e - g.b > (uint32_t)kMedium && g.e - g.b > T;
  return g;
}

// Emit of a pass: each workgroup takes kEmitBlock segments, counts per
// segment the next pass's segments, small segments, tiles, prims and split
// children, scans the counts within the workgroup, gets its offsets from
// the workgroups before it by decoupled look-back (one wave per count;
// workgroups only wait on lower ones, which the dispatcher starts first),
// and writes: S's node and its split children's nodes (numbered after all
// of the pass's children), the next pass's segments (with planes, tile
// offsets and the tile->segment map) and the small-segment list.  Halves
// of one-sided child splits get their bounds and planes from their prims.
// The last workgroup writes the pass totals and the next pass's bases and
// clears this pass's segment count (every workgroup has read it before it
// published; the last one has waited for all of them).  A pass with no
// segments only carries its bookkeeping forward.
constexpr int kEmitBlock = 512;
constexpr int kEmitCounts = 5;

__global__ __launch_bounds__(kEmitBlock) void k_emit2(Ctx c, int pass) {
  constexpr int NW = kEmitBlock / kWave;
  __shared__ uint32_t ss[kEmitCounts * (NW + 1)];
  __shared__ uint32_t sPre[kEmitCounts];
  __shared__ uint3 sFix[4 * kEmitBlock];  // (segment, first slot, size) of open grandchildren to reduce
  __shared__ uint2 sRev[2 * kEmitBlock];  // (first slot, size) of ranges to turn around
  __shared__ uint32_t sNumFix, sNumRev;
  __shared__ float sb[12 * NW];
  const int par = pass & 1, np = par ^ 1;
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
      c.levelStart[2 * pass + 1] = k->baseTop[par];
      c.levelStart[2 * pass + 2] = k->baseTop[par];
      c.levelItems[pass + 1] = 0;
    }
    return;
  }
  if (blockIdx.x * kEmitBlock >= n) return;
  const uint32_t nb = div_up(n, kEmitBlock);
  const uint32_t s = blockIdx.x * kEmitBlock + threadIdx.x;
  const uint32_t nodeBase = k->baseNode[par], smallBase = k->baseSmall[par], topBase = k->baseTop[par];
  const bool pts = k->allPoints != 0;
  const uint32_t tag = (uint32_t)pass + 1;
  const uint32_t T = c.T;
  uint32_t begin = 0, end = 0, mid = 0, mode = 0, split[2] = {0, 0}, nCL[2] = {0, 0};
  uint32_t cnt[kEmitCounts] = {0, 0, 0, 0, 0};  // open, small, tiles, prims, split children
  // everything the writes below read, loaded before the first store (which
  // the compiler would not move loads across) and before the scans
  int32_t ccb[12], gcb[24];
  Seg g0{};
  if (s < n) {
    const PairPlan &P = c.plan[s];
    begin = P.begin; end = P.end; mid = P.mid; mode = P.mode;
    split[0] = P.split[0]; split[1] = P.split[1];
    for (int j = 0; j < 12; j++) ccb[j] = P.ccb[j / 6][j % 6];
    nCL[0] = c.classTot[4 * (size_t)s + 0];
    nCL[1] = c.classTot[4 * (size_t)s + 2];
    g0 = c.segs[par][s];
    for (int j = 0; j < 24; j++) gcb[j] = acc_identity(j);
    if (split[0] | split[1]) {
      const int4 *a = reinterpret_cast<const int4 *>(c.accB + (size_t)s * kCbCopies * kAccB);
      const uint32_t cps = acc_copies(end - begin);
#pragma unroll
      for (uint32_t cp = 0; cp < kCbCopies; cp++) {
        if (cp >= cps) break;
#pragma unroll
        for (int q = 0; q < 6; q++) {
          const int4 x = a[cp * 6 + q];
          const int w = 4 * q;
          gcb[w] = (w % 6) < 3 ? min(gcb[w], x.x) : max(gcb[w], x.x);
          gcb[w + 1] = ((w + 1) % 6) < 3 ? min(gcb[w + 1], x.y) : max(gcb[w + 1], x.y);
          gcb[w + 2] = ((w + 2) % 6) < 3 ? min(gcb[w + 2], x.z) : max(gcb[w + 2], x.z);
          gcb[w + 3] = ((w + 3) % 6) < 3 ? min(gcb[w + 3], x.w) : max(gcb[w + 3], x.w);
        }
      }
    }
    for (int side = 0; side < 2; side++) {
      const uint32_t cb = side ? mid : begin, ce = side ? end : mid;
      if (!split[side]) { cnt[1]++; continue; }
      cnt[4]++;
      for (uint32_t gs = 0; gs < 2; gs++) {
        const GrandKid g = grand_kid(cb, ce, nCL[side], child_rev(mode, side), gs, T);
        if (g.large) { cnt[0]++; cnt[2] += div_up(g.e - g.b, kTile); cnt[3] += g.e - g.b; }
        else cnt[1]++;
      }
    }
  }
  if (threadIdx.x == 0) sNumFix = sNumRev = 0;
  uint32_t ex[kEmitCounts], tt[kEmitCounts];
  block_excl_scan_multi<kEmitBlock, kEmitCounts, true>(cnt, ex, tt, ss);
  const uint32_t w = threadIdx.x / kWave;
  if (w < (uint32_t)kEmitCounts) {
    uint32_t v = tt[0];
#pragma unroll
    for (int q = 1; q < kEmitCounts; q++) v = w == (uint32_t)q ? tt[q] : v;
    const uint32_t pre =
        lookback_exclusive_wave(c.emitStatus + (size_t)w * c.maxEmitBlocks, blockIdx.x, blockIdx.x, tag, v, false,
                                &c.ctr->spinFault);
    if (lane_id() == 0) sPre[w] = pre;
  }
  sync_block_lds();
  uint32_t pre[kEmitCounts];
#pragma unroll
  for (int q = 0; q < kEmitCounts; q++) pre[q] = sPre[q];
  if (blockIdx.x == nb - 1 && threadIdx.x == 0) {
    const uint32_t runL = pre[0] + tt[0], runS = pre[1] + tt[1], runT = pre[2] + tt[2], runI = pre[3] + tt[3];
    const uint32_t runX = pre[4] + tt[4];
    k->nodeCount = k->baseNode[np] = nodeBase + 2 * n + 2 * runX;
    k->numSmall = k->baseSmall[np] = smallBase + runS;
    k->numTop = k->baseTop[np] = topBase + n + runX;
    k->numLarge[np] = runL;
    k->numLarge[par] = 0;
    c.levelStart[2 * pass + 1] = topBase + n;
    c.levelStart[2 * pass + 2] = topBase + n + runX;
    k->numTiles = runT;
    k->ticket = 0;
    for (uint32_t q = 0; q < kTicketCounters; q++) c.tickets[kTicketStride * q] = 0;
    c.tileStart[np][runL] = runT;
    c.levelItems[pass + 1] = runI;
  }
  uint32_t fillStart[4] = {0, 0, 0, 0}, fillCount[4] = {0, 0, 0, 0}, fillSeg[4] = {0, 0, 0, 0};
  if (s < n) {
    uint32_t li = pre[0] + ex[0], si = pre[1] + ex[1], ti = pre[2] + ex[2], xi = pre[4] + ex[4];
    const Seg &g = g0;
    const uint32_t child = nodeBase + 2 * s;
    if (pts) {
      // point input: a node's box is its centroid bounds, final here
      Node nd;
      for (int d = 0; d < 3; d++) { nd.lo[d] = dec_f(g.cb[d]); nd.hi[d] = dec_f(g.cb[d + 3]); }
      nd.oc = pack_oc(child, 0);
      c.nodes[g.node] = nd;
    } else {
      c.nodes[g.node].oc = pack_oc(child, 0);
    }
    c.topNodes[topBase + s] = g.node;
    int nf = 0;
    for (int side = 0; side < 2; side++) {
      const uint32_t cb = side ? mid : begin, ce = side ? end : mid;
      const uint32_t rv = child_rev(mode, side);
      const uint32_t cid = child + side;
      if (!split[side]) {
        c.small[smallBase + si] = SmallSeg{cb, ce, cid, (uint32_t)np, cb, rv, 0, 0};
        si++;
        continue;
      }
      const uint32_t gid = nodeBase + 2 * n + 2 * xi;
      if (pts) {
        Node nd;
        for (int d = 0; d < 3; d++) { nd.lo[d] = dec_f(ccb[6 * side + d]); nd.hi[d] = dec_f(ccb[6 * side + d + 3]); }
        nd.oc = pack_oc(gid, 0);
        c.nodes[cid] = nd;
      } else {
        c.nodes[cid].oc = pack_oc(gid, 0);
      }
      c.topNodes[topBase + n + xi] = cid;
      xi++;
      if (rv && (nCL[side] == 0 || nCL[side] == ce - cb)) sRev[atomicAdd(&sNumRev, 1u)] = make_uint2(cb, ce - cb);
      for (uint32_t gs = 0; gs < 2; gs++) {
        const GrandKid gk = grand_kid(cb, ce, nCL[side], rv, gs, T);
        if (!gk.large) {
          c.small[smallBase + si] = SmallSeg{gk.b, gk.e, gid + gs, (uint32_t)np, gk.b, gk.rev, 0, 0};
          si++;
          continue;
        }
        Seg ns;
        ns.begin = gk.b; ns.end = gk.e; ns.node = gid + gs; ns.pbeg = gk.b; ns.rev = gk.rev; ns.pad = 0;
        for (int j = 0; j < 6; j++) ns.cb[j] = gcb[6 * (2 * side + gs) + j];
        c.segs[np][li] = ns;
        if (!gk.oneSided && gather_bounds(split[side], ce - cb)) {
          float lo[3], hi[3];
          for (int d = 0; d < 3; d++) { lo[d] = dec_f(ns.cb[d]); hi[d] = dec_f(ns.cb[d + 3]); }
          int dim;
          float pos;
          choose_split(lo, hi, dim, pos);
          c.segDim[np][li] = dim;
          c.segPos[np][li] = pos;
        } else {
          sFix[atomicAdd(&sNumFix, 1u)] = make_uint3(li, gk.b, gk.e - gk.b);
        }
        const uint32_t tiles = div_up(gk.e - gk.b, kTile);
        c.tileStart[np][li] = ti;
        fillStart[nf] = ti;
        fillCount[nf] = tiles;
        fillSeg[nf] = li;
        nf++;
        ti += tiles;
        li++;
      }
    }
  }
  // tile -> segment map: runs by their own lane, except the long runs of
  // the top passes' few big segments, which the whole wave writes
  constexpr uint32_t kOwnFill = 256;
  const uint32_t lane = lane_id();
  for (int f = 0; f < 4; f++) {
    if (fillCount[f] <= kOwnFill)
      for (uint32_t t = 0; t < fillCount[f]; t++) c.tileSeg[np][fillStart[f] + t] = fillSeg[f];
    uint64_t m = __ballot(fillCount[f] > kOwnFill);
    while (m) {
      const int src = __builtin_ctzll(m);
      m &= m - 1;
      const uint32_t st = __shfl(fillStart[f], src, kWave), cn = __shfl(fillCount[f], src, kWave);
      const uint32_t sg = __shfl(fillSeg[f], src, kWave);
      for (uint32_t t = lane; t < cn; t += kWave) c.tileSeg[np][st + t] = sg;
    }
  }
  // reversed children split one-sided: turn their ranges around (each
  // thread swaps a pair of slots), so their halves are stored in order
  __syncthreads();
  float4 *items = c.items[np];
  const uint32_t nRev = sNumRev;
  for (uint32_t f = 0; f < nRev; f++) {
    const uint32_t pb = sRev[f].x, sz = sRev[f].y;
    for (uint32_t i = threadIdx.x; i < sz / 2; i += kEmitBlock) {
      const float4 a = items[pb + i], b = items[pb + sz - 1 - i];
      items[pb + i] = b;
      items[pb + sz - 1 - i] = a;
    }
  }
  // open grandchildren whose bounds pass B did not gather (halves of
  // one-sided child splits, children of <= kLazyBounds prims): centroid
  // bounds and plane from their prims (a full barrier: the segment records
  // above must land before these overwrite their bounds, and the turned
  // ranges before these read them)
  __syncthreads();
  const uint32_t nFix = sNumFix;
  for (uint32_t f = 0; f < nFix; f++) {
    const uint32_t li = sFix[f].x, pb = sFix[f].y, sz = sFix[f].z;
    float v[12];
    for (int j = 0; j < 12; j++) v[j] = (j % 6) < 3 ? INFINITY : -INFINITY;
    for (uint32_t i = threadIdx.x; i < sz; i += kEmitBlock) {
      const float4 it = items[pb + i];
      v[0] = fminf(v[0], it.x); v[1] = fminf(v[1], it.y); v[2] = fminf(v[2], it.z);
      v[3] = fmaxf(v[3], it.x); v[4] = fmaxf(v[4], it.y); v[5] = fmaxf(v[5], it.z);
    }
    block_reduce_bounds<kEmitBlock>(v, sb);
    if (threadIdx.x < 6) c.segs[np][li].cb[threadIdx.x] = enc_f(v[threadIdx.x]);
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
// cond[k] says whether step k's source lane lies in the lane's segment (and,
// for the broadcast steps, that the lane's row takes part in the step).
// All scans run as unsigned max scans: a min becomes a max of complemented
// keys, so 0 is the identity of every step, which is what a row_shr step
// reads from a lane outside the row (bound_ctrl), and no step needs its
// destination preloaded with an identity.
template <bool MIN>
__device__ __forceinline__ uint32_t seg_key(int32_t v) {
  const uint32_t u = (uint32_t)v ^ 0x80000000u;  // signed order -> unsigned order
  return MIN ? ~u : u;
}
template <bool MIN>
__device__ __forceinline__ int32_t seg_unkey(uint32_t k) {
  return (int32_t)((MIN ? ~k : k) ^ 0x80000000u);
}
__device__ __forceinline__ uint32_t seg_scan_max(uint32_t v, const bool (&cond)[6]) {
#define CUBQL_SEG_STEP(K, CTRL, ROWMASK, BOUND)                                                    \
  {                                                                                                \
    const uint32_t t = (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, ROWMASK, 0xF, BOUND);      \
    const uint32_t m = max(v, t);                                                                  \
    v = cond[K] ? m : v;                                                                           \
  }
  CUBQL_SEG_STEP(0, 0x111, 0xF, true)
  CUBQL_SEG_STEP(1, 0x112, 0xF, true)
  CUBQL_SEG_STEP(2, 0x114, 0xF, true)
  CUBQL_SEG_STEP(3, 0x118, 0xF, true)
  CUBQL_SEG_STEP(4, 0x142, 0xA, false)
  CUBQL_SEG_STEP(5, 0x143, 0xC, false)
#undef CUBQL_SEG_STEP
  return v;
}

// Reduction of v over the lane's segment [sb, se), in every lane of it.
template <bool MIN>
__device__ __forceinline__ int32_t seg_reduce(int32_t v, const bool (&cond)[6], uint32_t last) {
  return seg_unkey<MIN>((uint32_t)__shfl((int)seg_scan_max(seg_key<MIN>(v), cond), (int)last, kWave));
}

// Inclusive unsigned max scan over the whole wave (0 is the identity).
__device__ __forceinline__ uint32_t wave_scan_max_u(uint32_t v) {
#define CUBQL_MAX_STEP(CTRL, ROWMASK)                                                               \
  {                                                                                                 \
    const uint32_t t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWMASK, 0xF, true);  \
    v = max(v, t);                                                                                  \
  }
  CUBQL_MAX_STEP(0x111, 0xF)
  CUBQL_MAX_STEP(0x112, 0xF)
  CUBQL_MAX_STEP(0x114, 0xF)
  CUBQL_MAX_STEP(0x118, 0xF)
  CUBQL_MAX_STEP(0x142, 0xA)
  CUBQL_MAX_STEP(0x143, 0xC)
#undef CUBQL_MAX_STEP
  return v;
}

// Segmented max of 26-bit values by an unsegmented scan: the key puts the
// lane's segment head above the value, heads grow along the lanes, so the
// scan's maximum over lanes [0, l] comes from l's own segment.
constexpr uint32_t kRelBits = 26, kRelMask = (1u << kRelBits) - 1u;
__device__ __forceinline__ uint32_t rel_seg_max(uint32_t sbHi, uint32_t u, uint32_t last) {
  return (uint32_t)__shfl((int)wave_scan_max_u(sbHi | u), (int)last, kWave) & kRelMask;
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
// that holds its position; bounds are segmented wave reductions, ranks are
// ballot prefix counts less those of the segment's head, items move with
// one ds_permute per value.  Split rules and partition order (left stable, right reversed,
// one-sided -> middle) are those of the CPU recursion
// (cpu/spatialMedian.h:81-114), so positions and leaves are identical to a
// node-at-a-time build; local ids are assigned breadth-first.
template <bool PTS>
__device__ void reg_subtree(const Ctx &c, uint32_t pos0, uint32_t cnt, float x, float y, float z,
                            uint32_t id, float bx0, float by0, float bz0, float bx1, float by1,
                            float bz1, uint32_t rootLocal, uint32_t &nLocal, Node *loc) {
  const uint32_t lane = lane_id();
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
  bool first = true;
  // point input whose root extent spans < 2^26 encoded floats per axis:
  // the segmented reductions of later rounds run as unsegmented scans on
  // (segment head, value - root minimum) keys
  bool rel = false, relPos = false;
  int32_t eLo[3] = {0, 0, 0};
  while (__ballot(active)) {
    float clo[3], chi[3], nlo[3], nhi[3];
    if (PTS && rel) {
      const uint32_t sbHi = sb << kRelBits, last = se - 1;
      const int32_t e3[3] = {xe, ye, ze};
      int32_t lo[3], hi[3];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        const uint32_t u = (uint32_t)(e3[d] - eLo[d]);
        lo[d] = eLo[d] + (int32_t)(rel_seg_max(sbHi, u ^ kRelMask, last) ^ kRelMask);
        hi[d] = eLo[d] + (int32_t)rel_seg_max(sbHi, u, last);
      }
      // (the encoding of a non-negative float is its bit pattern)
      if (relPos) {
        for (int d = 0; d < 3; d++) { clo[d] = __int_as_float(lo[d]); chi[d] = __int_as_float(hi[d]); }
      } else {
        for (int d = 0; d < 3; d++) { clo[d] = dec_f(lo[d]); chi[d] = dec_f(hi[d]); }
      }
    } else if (first) {
      // one segment, lanes [0, cnt): plain wave reductions, idle lanes
      // holding the identity
      first = false;
      clo[0] = dec_f(wave_min_i32(active ? xe : INT32_MAX));
      clo[1] = dec_f(wave_min_i32(active ? ye : INT32_MAX));
      clo[2] = dec_f(wave_min_i32(active ? ze : INT32_MAX));
      chi[0] = dec_f(wave_max_i32(active ? xe : INT32_MIN));
      chi[1] = dec_f(wave_max_i32(active ? ye : INT32_MIN));
      chi[2] = dec_f(wave_max_i32(active ? ze : INT32_MIN));
      if (!PTS) {
        nlo[0] = dec_f(wave_min_i32(active ? b0 : INT32_MAX));
        nlo[1] = dec_f(wave_min_i32(active ? b1 : INT32_MAX));
        nlo[2] = dec_f(wave_min_i32(active ? b2 : INT32_MAX));
        nhi[0] = dec_f(wave_max_i32(active ? b3 : INT32_MIN));
        nhi[1] = dec_f(wave_max_i32(active ? b4 : INT32_MIN));
        nhi[2] = dec_f(wave_max_i32(active ? b5 : INT32_MIN));
      }
      if (PTS) {
        rel = true;
        relPos = true;
#pragma unroll
        for (int d = 0; d < 3; d++) {
          eLo[d] = enc_cmp(clo[d]);
          rel = rel && (int64_t)enc_cmp(chi[d]) - eLo[d] <= (int64_t)kRelMask;
          relPos = relPos && eLo[d] >= 0;
        }
        // lanes past the prims hold the minimum (relative value 0)
        if (lane >= cnt) { xe = eLo[0]; ye = eLo[1]; ze = eLo[2]; }
      }
    } else {
      const int h = (int)sb, l = (int)lane;
      // (the broadcast steps write rows 1, 3 and rows 2, 3 only)
      const bool cond[6] = {l - 1 >= h, l - 2 >= h, l - 4 >= h, l - 8 >= h,
                            (lane & 16u) != 0 && (int)(lane & ~15u) - 1 >= h, lane >= 32u && 31 >= h};
      const uint32_t last = se - 1;
      clo[0] = dec_f(seg_reduce<true>(xe, cond, last));
      clo[1] = dec_f(seg_reduce<true>(ye, cond, last));
      clo[2] = dec_f(seg_reduce<true>(ze, cond, last));
      chi[0] = dec_f(seg_reduce<false>(xe, cond, last));
      chi[1] = dec_f(seg_reduce<false>(ye, cond, last));
      chi[2] = dec_f(seg_reduce<false>(ze, cond, last));
      if (!PTS) {
        nlo[0] = dec_f(seg_reduce<true>(b0, cond, last));
        nlo[1] = dec_f(seg_reduce<true>(b1, cond, last));
        nlo[2] = dec_f(seg_reduce<true>(b2, cond, last));
        nhi[0] = dec_f(seg_reduce<false>(b3, cond, last));
        nhi[1] = dec_f(seg_reduce<false>(b4, cond, last));
        nhi[2] = dec_f(seg_reduce<false>(b5, cond, last));
      }
    }
    if (PTS)
      for (int d = 0; d < 3; d++) { nlo[d] = clo[d]; nhi[d] = chi[d]; }
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
    const bool isL = split && ve < (relPos ? __float_as_int(pos) : enc_cmp(pos));
    // ranks inside the segment: counts below the lane minus the counts
    // below its head, read from the head with the children's local id
    // (7 + 7 + 18 bits); the left count of the segment from its last lane
    const uint64_t mL = __ballot(isL), mR = __ballot(split && !isL);
    const uint64_t heads = __ballot(split && isHead);
    const uint32_t pL = mbcnt(mL), pR = mbcnt(mR);
    const uint32_t hw = (uint32_t)__shfl((int)(pL | (pR << 7) | ((nLocal + 2u * mbcnt(heads)) << 14)), (int)sb, kWave);
    const uint32_t tL = (uint32_t)__shfl((int)(pL + (isL ? 1u : 0u)), (int)(se - 1u), kWave);
    const uint32_t hL = hw & 127u, hR = (hw >> 7) & 127u, child = hw >> 14;
    const uint32_t nl = tL - hL;
    const uint32_t dst = !split ? lane : (isL ? sb + (pL - hL) : se - 1u - (pR - hR));
    // (one-sided: the middle of the global range, spatialMedian.h:107;
    // (2 pos0 + sb + se) / 2 - pos0 == (sb + se) / 2)
    const uint32_t mid = (nl != 0 && nl != n) ? sb + nl : (sb + se) >> 1;
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

// Finishes the node of range [gb, gb + m) (m <= kSmall, or a leaf of any
// size up to T) whose local id is rootMe, numbering new nodes from nLocal
// into loc.  Its items are gSrc[0 .. m) in the node's order, or (rev) in
// reverse.
template <bool PTS>
__device__ void small_sub(const Ctx &c, const float4 *gSrc, bool rev, uint32_t gb, uint32_t m, uint32_t rootMe,
                          uint32_t &nLocal, Node *loc, float4 *sIt, uint32_t *sFr) {
  const uint32_t lane = lane_id();
  const uint32_t T = c.T;
  auto ld = [&](uint32_t i) { return gSrc[rev ? m - 1 - i : i]; };
  if (m <= T && m > 64) {
    // leaf of any size up to T
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = lane; i < m; i += kWave) {
      float4 v = ld(i);
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
    float4 v = lane < m ? ld(lane) : make_float4(0.f, 0.f, 0.f, 0.f);
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
        if (i < m) t[k] = ld(i);
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
  uint32_t b, e, me, buf;  // buf: bit 0 buffer, bit 1 bounds unknown, bit 2 reversed
  uint32_t pb, pad[3];     // first slot of the piece's items
  int32_t cb[6];
  int32_t pad2[2];
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
  // pieces written by this wave sit in their own range, in order
  auto piece = [&](uint32_t buf, uint32_t b, uint32_t n, uint32_t me) {
    small_sub<PTS>(c, c.items[buf] + gb + b, false, gb + b, n, me, nLocal, loc, sIt, sFr);
    sync_wave_lds();
  };
  if (m <= (uint32_t)kSmall || m <= T) {
    small_sub<PTS>(c, c.items[sg.buf] + sg.pbeg, sg.rev != 0, gb, m, 0u, nLocal, loc, sIt, sFr);
  } else {
    int sp = 1;
    bool known = false;  // bounds of the popped frame
    if (lane == 0) {
      MedFrame f0{};
      f0.b = 0u; f0.e = m; f0.me = 0u; f0.buf = sg.buf | 2u | (sg.rev ? 4u : 0u);
      f0.pb = sg.pbeg;
      sMed[0] = f0;
    }
    sync_wave_lds();
    while (sp > 0) {
      --sp;
      const MedFrame f = sMed[sp];
      const uint32_t b = f.b, e = f.e, me = f.me, bf = f.buf & 1u;
      known = (f.buf & 2u) == 0;
      const bool rev = (f.buf & 4u) != 0;
      const uint32_t pb = f.pb;
      const float4 *src = c.items[bf];
      float4 *dst = c.items[bf ^ 1u];
      int32_t cb[6], bb[6];
      if (known) {
        for (int j = 0; j < 6; j++) cb[j] = bb[j] = f.cb[j];
      } else {
        // (bounds do not depend on the order)
        med_bounds<PTS>(c, src, pb, pb + (e - b), cb, bb);
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
      // (only a piece of more than 2 kSmall prims can have a child that
      // needs them; a child of another piece that is still larger than
      // kSmall after an uneven split takes a bounds pass)
      const bool gather = PTS && e - b > 2u * (uint32_t)kSmall;
      uint32_t nL = 0, nR = 0;
      for (uint32_t c0 = b; c0 < e; c0 += kMedChunk * kWave) {
        float4 v[kMedChunk];
        bool act[kMedChunk];
#pragma unroll
        for (int k = 0; k < kMedChunk; k++) {
          const uint32_t i = c0 + (uint32_t)k * kWave + lane;
          act[k] = i < e;
          v[k] = act[k] ? src[rev ? pb + (e - 1 - i) : pb + (i - b)] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int k = 0; k < kMedChunk; k++) {
          const int32_t ec = enc_cmp(dim == 0 ? v[k].x : (dim == 1 ? v[k].y : v[k].z));
          const bool isL = act[k] && ec < posE;
          const uint64_t mL = __ballot(isL);
          const uint64_t mR = __ballot(act[k] && !isL);
          if (act[k]) {
            const uint32_t d = isL ? b + nL + mbcnt(mL) : e - 1 - (nR + mbcnt(mR));
            dst[gb + d] = v[k];
            if (gather) {
              const int32_t ex = enc_cmp(v[k].x), ey = enc_cmp(v[k].y), ez = enc_cmp(v[k].z);
              const int s = isL ? 0 : 1;
              // (select both sides' updates: no divergent array indexing)
              for (int q = 0; q < 2; q++)
                if (s == q) bounds_add(sc[q], ex, ey, ez, ex, ey, ez);
            }
          }
          nL += (uint32_t)__popcll(mL);
          nR += (uint32_t)__popcll(mR);
        }
      }
      const bool oneSided = nL == 0 || nR == 0;
      // (one-sided: the middle of the global range, spatialMedian.h:107;
      // (2 gb + b + e) / 2 - gb == (b + e) / 2)
      const uint32_t mid = !oneSided ? b + nL : (b + e) >> 1;
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
          const bool carry = gather && !oneSided;
          g.b = cb0; g.e = ce0; g.me = child + side; g.buf = (bf ^ 1u) | (carry ? 0u : 2u);
          g.pb = gb + cb0;
          if (carry) {
            // (wave-uniform branch: every lane joins the reductions)
            int32_t rc[6];
            for (int j = 0; j < 6; j++) rc[j] = side ? sc[1][j] : sc[0][j];
            bounds_wave(rc);
            for (int j = 0; j < 6; j++) g.cb[j] = rc[j];
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

// (pp: parity of the pass after the last one launched; while segments are
// still open there the small phase does nothing -- the host launches it
// speculatively after the passes it expects the tree to need)
__global__ __launch_bounds__(kWave) void k_small(Ctx c, int pp) {
  __shared__ float4 sIt[kSmall];
  __shared__ uint32_t sFr[kSmall + 2];
  __shared__ MedFrame sMed[kMedium / (kSmall + 1) + 1];
  // (launched before the host has checked the passes: after a look-back
  // wait past its bound the items are not trusted, so nothing is gathered)
  if (c.ctr->numLarge[pp] != 0 || c.ctr->spinFault != 0) return;
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
__global__ __launch_bounds__(1024) void k_small_scan(Ctx c, uint32_t tag, int pp) {
  __shared__ uint32_t ss[1024 / kWave + 2];
  if (c.ctr->numLarge[pp] != 0) return;
  count_scan_tile(c.smallCount, c.smallBase, c.ctr->numSmall, c.ctr->nodeCount, c.tileStatus, tag,
                  &c.ctr->numNodesOut, ss, 1u, &c.ctr->spinFault);
}
// (one wave per segment, four segments a workgroup)
__global__ __launch_bounds__(256) void k_small_emit(Ctx c, Node *out) {
  const uint32_t nSmall = c.ctr->numSmall;
  const uint32_t wpb = blockDim.x / kWave;
  // (the segment index is wave-uniform: its fields come in scalar loads)
  const uint32_t w0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * wpb + threadIdx.x / kWave));
  for (uint32_t w = w0; w < nSmall; w += gridDim.x * wpb) {
    const SmallSeg sg = c.small[w];
    copy_local_nodes(c.loc + 2 * (size_t)sg.begin, c.smallCount[w], sg.node, c.smallBase[w], out);
  }
}

__global__ void k_copy_top(const Node *src, Node *dst, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// Bounds of the pass-phase inner nodes, node list by node list (two a
// pass: a pass's split children before its segments), deepest first.
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
  ScratchArena scr(mem, s);
  Ctx c{};
  c.boxes = boxes;
  c.n = n;
  c.T = T;
  const uint32_t numBlocks = div_up(n, kTile);
  c.maxLarge = n / (kMedium + 1) + 2;
  const uint32_t maxTiles = div_up(n, kTile) + c.maxLarge + 1;
  const uint32_t maxPasses = 2048;
  CUBQL_CHECK(scr.reserve(&c.items[0], n));
  CUBQL_CHECK(scr.reserve(&c.items[1], n));
  CUBQL_CHECK(scr.reserve(&c.nodes, 2 * (size_t)n + 2));
  CUBQL_CHECK(scr.reserve(&c.ctr, 1));
  CUBQL_CHECK(scr.reserve(&c.blockCount, numBlocks + 1));
  CUBQL_CHECK(scr.reserve(&c.segs[0], c.maxLarge));
  CUBQL_CHECK(scr.reserve(&c.segs[1], c.maxLarge));
  const uint32_t maxEmitBlocks = div_up(c.maxLarge, kEmitBlock);
  CUBQL_CHECK(scr.reserve(&c.accA, (size_t)kAccA * c.maxLarge * kCbCopies));
  CUBQL_CHECK(scr.reserve(&c.accB, (size_t)kAccB * c.maxLarge * kCbCopies));
  CUBQL_CHECK(scr.reserve(&c.plan, c.maxLarge));
  CUBQL_CHECK(scr.reserve(&c.classTot, 4 * (size_t)c.maxLarge));
  CUBQL_CHECK(scr.reserve(&c.tileLeft, maxTiles));
  CUBQL_CHECK(scr.reserve(&c.tilePre, maxTiles));
  for (int p = 0; p < 2; p++) {
    CUBQL_CHECK(scr.reserve(&c.tileStart[p], c.maxLarge + 1));
    CUBQL_CHECK(scr.reserve(&c.tileSeg[p], maxTiles));
    CUBQL_CHECK(scr.reserve(&c.segDim[p], c.maxLarge));
    CUBQL_CHECK(scr.reserve(&c.segPos[p], c.maxLarge));
  }
  c.maxEmitBlocks = maxEmitBlocks;
  CUBQL_CHECK(scr.reserve(&c.emitStatus, (size_t)kEmitCounts * maxEmitBlocks));
  CUBQL_CHECK(scr.reserve(&c.tickets, kTicketStride * kTicketCounters));
  c.maxTiles = maxTiles;
  CUBQL_CHECK(scr.reserve(&c.tileStatus, 4 * (size_t)maxTiles));
  CUBQL_CHECK(scr.reserve(&c.small, n + 1));
  CUBQL_CHECK(scr.reserve(&c.topNodes, n + 1));
  CUBQL_CHECK(scr.reserve(&c.levelStart, 2 * (size_t)maxPasses + 2));
  CUBQL_CHECK(scr.reserve(&c.levelItems, (size_t)maxPasses + 2));
  CUBQL_CHECK(scr.reserve(&c.blockBounds, 8 * (size_t)numBlocks + 8));
  CUBQL_CHECK(scr.reserve(&c.loc, 2 * (size_t)n + 64));
  CUBQL_CHECK(scr.reserve(&c.smallCount, (size_t)n + 1));
  CUBQL_CHECK(scr.reserve(&c.smallBase, (size_t)n + 1));
  CUBQL_CHECK(scr.commit());

  uint32_t *primIDs = nullptr;
  CUBQL_CHECK(mem_alloc(mem, (void **)&primIDs, (size_t)n * sizeof(uint32_t), s));
  c.primIDs = primIDs;
  auto fail = [&](int st) {
    mem_free(mem, primIDs, s);
    return st;
  };

  const size_t nStatus = 8 * (size_t)maxTiles, nEmit = 2 * (size_t)kEmitCounts * maxEmitBlocks;
  const size_t initWork = std::max<size_t>({(size_t)c.maxLarge * (kCbCopies * kAccA / 4), nStatus, nEmit});
  k_init_acc<<<std::max<uint32_t>(1u, std::min<uint32_t>(div_up(initWork, 256), 2048)), 256, 0, s>>>(c, nStatus,
                                                                                                   nEmit);
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
  // (pass B is persistent: one resident wave of workgroups)
  int partPerCU = 0, boundsPerCU = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&partPerCU, k_pair, kBlock, 0) != hipSuccess || partPerCU < 1)
    partPerCU = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&boundsPerCU, k_bounds, kBlock, 0) != hipSuccess || boundsPerCU < 1)
    boundsPerCU = 1;
  const uint32_t partGrid =
      std::max<uint32_t>(kTicketCounters, std::min<uint32_t>(maxTiles, (uint32_t)numCU * (uint32_t)partPerCU));
  const uint32_t boundsGrid = std::min<uint32_t>(maxTiles, (uint32_t)numCU * (uint32_t)boundsPerCU);
  const uint32_t planGrid = div_up(c.maxLarge, 256 / kWave);
  const uint32_t scanBlocks = div_up(maxTiles, kTileScanPart);
  int pass = 0;
  std::vector<std::pair<int, int>> partSlots, boundsSlots;  // (profile slot, pass)
  // The first batch covers the depth a well-spread input needs (halving
  // down to kMedium prims, two levels a pass): a pass past the tree's depth
  // (four near-empty launches) costs about what a host check does, ~20 us,
  // so deeper trees take a check and another batch instead of slack here.
  int levels = 0;
  for (uint64_t m = (uint64_t)kMedium; m < n; m *= 2) levels++;
  int batch = std::max(kPassesPerSync, (levels + 1) / 2);
  bool firstBatch = true;
  int specPass = -1, smallSlot = -1;
  uint32_t specGrid = 0;
  while (true) {
    for (int i = 0; i < batch && pass < (int)maxPasses; i++, pass++) {
      int bs = prof_begin(s);
      k_bounds<<<boundsGrid, kBlock, 0, s>>>(c, pass);
      prof_end(bs, s, "sm_bounds");
      if (bs >= 0) boundsSlots.push_back({bs, pass});
      k_plan<<<scanBlocks + planGrid, 256, 0, s>>>(c, pass, scanBlocks);
      int ps = prof_begin(s);
      k_pair<<<partGrid, kBlock, 0, s>>>(c, pass);
      prof_end(ps, s, "sm_partition");
      if (ps >= 0) partSlots.push_back({ps, pass});
      k_emit2<<<maxEmitBlocks, kEmitBlock, 0, s>>>(c, pass);
    }
    if (firstBatch) {
      // The small phase, launched before the host knows the passes are
      // done (it does nothing otherwise): one workgroup per segment needs
      // the segment count, and after P passes there are at most
      // 4 (4^P - 1) / 3 + 1 small segments (each open segment of a pass
      // makes at most 4 open and 4 small ones).
      uint64_t bound = 1, open = 1;
      for (int p = 0; p < pass && bound <= (uint64_t)n; p++) {
        bound += 4 * open;
        open *= 4;
      }
      specGrid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(bound, (uint64_t)n + 1));
      smallSlot = prof_begin(s);
      k_small<<<specGrid, kWave, 0, s>>>(c, pass & 1);
      prof_end(smallSlot, s, "sm_small");
      // (status tag beyond every pass's)
      k_small_scan<<<div_up(specGrid, kCountScanTile), 1024, 0, s>>>(c, maxPasses + 2, pass & 1);
      specPass = pass;
      firstBatch = false;
    }
    if (hipGetLastError() != hipSuccess) return fail((set_error("pass launch failed"), CUBQL_ERR_HIP));
    if (hipMemcpyAsync(hw, c.ctr, sizeof(Counters), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return fail((set_error("pass sync failed"), CUBQL_ERR_HIP));
    const Counters *hc = (const Counters *)hw;
    if (hc->spinFault)
      return fail((set_error("spatial-median build: a look-back wait exceeded its bound"), CUBQL_ERR_INTERNAL));
    // open segments of the next pass to run
    if (hc->numLarge[pass & 1] == 0) break;
    batch = kPassesPerSync;
    if (pass >= (int)maxPasses) return fail((set_error("tree deeper than %u levels", 2 * maxPasses), CUBQL_ERR_INTERNAL));
  }
  const Counters hc0 = *(const Counters *)hw;
  const uint32_t numSmall = hc0.numSmall;
  const uint32_t numValid = hc0.numValid;
  const uint32_t numTop = hc0.numTop;
  const uint32_t topCount = hc0.nodeCount;  // ids [0, topCount) belong to the pass phase
  uint32_t nodesOut = hc0.numNodesOut;
  if (specPass != pass) {
    // the speculative small phase found open segments: run it now, one
    // workgroup per segment (the dispatcher balances the uneven segment
    // costs far better than a fixed round-robin over resident waves)
    smallSlot = prof_begin(s);
    k_small<<<std::max<uint32_t>(1u, numSmall), kWave, 0, s>>>(c, pass & 1);
    prof_end(smallSlot, s, "sm_small");
    k_small_scan<<<std::max<uint32_t>(1u, div_up(numSmall, kCountScanTile)), 1024, 0, s>>>(c, maxPasses + 2, pass & 1);
    if (hipGetLastError() != hipSuccess) return fail((set_error("small-phase launch failed"), CUBQL_ERR_HIP));
    if (hipMemcpyAsync(hw, c.ctr, sizeof(Counters), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return fail((set_error("final sync failed"), CUBQL_ERR_HIP));
    if (((const Counters *)hw)->spinFault)
      return fail((set_error("spatial-median build: a look-back wait exceeded its bound"), CUBQL_ERR_INTERNAL));
    nodesOut = ((const Counters *)hw)->numNodesOut;
  }
  const uint32_t numNodes = numValid ? (numSmall ? nodesOut : topCount) : 0;
  if (prof_on()) {
    // algorithmic bytes: pass B reads and writes one 16-byte item per prim
    // of the pass, pass A reads it; the small phase reads each item once,
    // writes its primID and the nodes it creates (gathering boxes for
    // non-point input)
    std::vector<uint32_t> li(pass + 1, 0);
    if (hipMemcpy(li.data(), c.levelItems, sizeof(uint32_t) * pass, hipMemcpyDeviceToHost) == hipSuccess) {
      for (auto &ps : partSlots) prof_set(ps.first, 32.0 * li[ps.second], li[ps.second]);
      for (auto &bs : boundsSlots) prof_set(bs.first, 16.0 * li[bs.second], li[bs.second]);
    }
    const bool pts = hc0.allPoints != 0;
    const double smallNodes = (double)numNodes - (double)topCount + (double)numSmall;
    prof_set(smallSlot, (16.0 + 4.0 + (pts ? 0.0 : 24.0)) * numValid + 32.0 * smallNodes, numValid);
    prof_flush();
  }
  Node *nodes = nullptr;
  const int st = mem_alloc(mem, (void **)&nodes, std::max<size_t>(numNodes, 2) * sizeof(Node), s);
  if (st) return fail(st);
  if (numNodes) {
    k_copy_top<<<std::max<uint32_t>(1u, std::min<uint32_t>(div_up(topCount, 256), 1024)), 256, 0, s>>>(
        c.nodes, nodes, topCount);
    if (numSmall) k_small_emit<<<div_up(numSmall, 4), 256, 0, s>>>(c, nodes);
    // (point input: the emit already wrote the pass-phase boxes; two node
    // lists a pass, each refit after the deeper ones)
    if (numTop && !hc0.allPoints)
      for (int l = 2 * pass - 1; l >= 0; l--) k_refit_top<<<64, 256, 0, s>>>(c, nodes, l);
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
