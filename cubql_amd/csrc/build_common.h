// build_common.h -- device helpers shared by the builders.
#pragma once
#include "common.h"

namespace cubql {

__device__ inline bool box_is_valid(const cubql_box3f &b) { return b.lower[0] <= b.upper[0]; }

__device__ inline cubql_box3f load_box(const cubql_box3f *p, uint32_t i) {
  const float2 *f = reinterpret_cast<const float2 *>(p + i);
  float2 a = f[0], b = f[1], c = f[2];
  cubql_box3f r;
  r.lower[0] = a.x; r.lower[1] = a.y; r.lower[2] = b.x;
  r.upper[0] = b.y; r.upper[1] = c.x; r.upper[2] = c.y;
  return r;
}

__device__ inline float coord(const float4 &v, int d) {
  return d == 0 ? v.x : (d == 1 ? v.y : v.z);
}

// Split choice (sm_builder.h:194-220 / spatialMedian.h:90-91): widest axis
// by strict '>' (ties to the lower axis), pos = midpoint.
__device__ inline void choose_split(const float lo[3], const float hi[3], int &dim, float &pos) {
  float s0 = hi[0] - lo[0], s1 = hi[1] - lo[1], s2 = hi[2] - lo[2];
  const bool d2 = s2 > fmaxf(s0, s1), d1 = !d2 && s1 > s0;
  dim = d2 ? 2 : (d1 ? 1 : 0);
  // (selects, not a branch on dim)
  const float l = d2 ? lo[2] : (d1 ? lo[1] : lo[0]);
  const float h = d2 ? hi[2] : (d1 ? hi[1] : hi[0]);
  pos = (l + h) / 2.f;
}

// One wave copies a finished small segment's cnt local nodes to the final
// array: local node 0 to `root`, local node j >= 1 to base + j - 1, with an
// inner node's child offset (a local id) rebased the same way.  A node moves
// as two 16-byte words; the second holds hi[1], hi[2] and oc.
__device__ __forceinline__ void copy_local_nodes(const Node *loc, uint32_t cnt, uint32_t root, uint32_t base,
                                                 Node *out) {
  const uint4 *src = reinterpret_cast<const uint4 *>(loc);
  for (uint32_t j = lane_id(); j < cnt; j += kWave) {
    const uint4 a = src[2 * j];
    uint4 b = src[2 * j + 1];
    const uint64_t oc = ((uint64_t)b.w << 32) | b.z;
    if (oc_count(oc) == 0) {
      const uint64_t g = pack_oc(base + (uint32_t)oc_offset(oc) - 1, 0);
      b.z = (uint32_t)g;
      b.w = (uint32_t)(g >> 32);
    }
    uint4 *d = reinterpret_cast<uint4 *>(out + (j == 0 ? root : base + j - 1));
    d[0] = a;
    d[1] = b;
  }
}

// One wave writes a 32-byte node: lanes 0-7 store one word each.
__device__ __forceinline__ void store_node_wave(Node *dst, const float (&lo)[3], const float (&hi)[3],
                                                uint64_t oc) {
  const uint32_t l = lane_id();
  uint32_t w = 0;
  w = (l == 0) ? __float_as_uint(lo[0]) : w;
  w = (l == 1) ? __float_as_uint(lo[1]) : w;
  w = (l == 2) ? __float_as_uint(lo[2]) : w;
  w = (l == 3) ? __float_as_uint(hi[0]) : w;
  w = (l == 4) ? __float_as_uint(hi[1]) : w;
  w = (l == 5) ? __float_as_uint(hi[2]) : w;
  w = (l == 6) ? (uint32_t)oc : w;
  w = (l == 7) ? (uint32_t)(oc >> 32) : w;
  if (l < 8) reinterpret_cast<uint32_t *>(dst)[l] = w;
}

// Tile status word for decoupled look-back: level tag (bits 34+), flag
// (bits 32-33: 0 empty, 1 local count, 2 inclusive prefix), value.
__device__ inline uint64_t status_word(uint32_t tag, uint32_t flag, uint32_t val) {
  return ((uint64_t)tag << 34) | ((uint64_t)flag << 32) | val;
}

// First half of the look-back below, which a caller may issue early: tile
// t's own count, as its inclusive prefix if it is the node's first tile
// (k == 0), else as an aggregate.
__device__ inline void lookback_publish(uint64_t *status, uint32_t t, uint32_t k, uint32_t tag, uint32_t local) {
  if (lane_id() == 0)
    __hip_atomic_store(&status[t], status_word(tag, k == 0 ? 2u : 1u, local), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// Decoupled look-back, wave-cooperative: all 64 lanes of one wave call it
// for tile t, the k-th tile of its node.  Publishes `local` (unless the
// caller already did: published), then reads the status words of up to 64
// preceding tiles at once (lane i reads tile t-1-i), sums them up to the
// nearest tile whose inclusive prefix is known and repeats further back if
// none is.  Tiles are claimed in ticket order, so every tile waited on is
// held by a running workgroup that publishes without waiting on later
// tiles.  Returns the node's sum over its earlier tiles (in every lane).
// Bound on the polls of one look-back (each one a status load and a short
// sleep, ~1 s in all).  Every tile waited on is held by a running
// workgroup, so a legitimate wait lasts microseconds; a wait past the bound
// means a tile was never published.  It then sets *err (when given) and
// goes on with the counts already seen: the prefix is at most the true one,
// so the caller's writes stay inside its node's range, and the host, which
// reads *err at its next synchronisation, fails the build instead of
// hanging the queue.
constexpr uint32_t kSpinLimit = 1u << 20;

__device__ inline uint32_t lookback_exclusive_wave(uint64_t *status, uint32_t t, uint32_t k, uint32_t tag,
                                                   uint32_t local, bool published = false,
                                                   uint32_t *err = nullptr) {
  const uint32_t lane = lane_id();
  if (!published) lookback_publish(status, t, k, tag, local);
  if (k == 0) return 0;
  uint32_t prefix = 0;
  uint32_t first = t - 1;  // tile read by lane 0
  uint32_t remaining = k;  // earlier tiles of this node: t-k .. t-1
  uint32_t spins = 0;
  while (true) {
    const bool inRange = lane < remaining;
    uint64_t w = 0;
    if (inRange) w = __hip_atomic_load(&status[first - lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t flag = ((uint32_t)(w >> 34) == tag) ? (uint32_t)(w >> 32) & 3u : 0u;
    const uint64_t incl = __ballot(inRange && flag == 2);
    const uint64_t notReady = __ballot(inRange && flag == 0);
    const uint32_t stop = incl ? (uint32_t)__builtin_ctzll(incl) : 63u;  // lanes 0..stop count
    const uint64_t upto = stop >= 63 ? ~0ull : ((1ull << (stop + 1)) - 1);
    const bool timedOut = ++spins > kSpinLimit;
    if ((notReady & upto) && !timedOut) {
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    uint32_t v = (inRange && lane <= stop && flag != 0) ? (uint32_t)w : 0u;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    prefix += v;
    if (timedOut && (notReady & upto)) {
      if (lane == 0 && err) atomicOr(err, 1u);
      break;
    }
    if (incl) break;
    first -= 64;
    remaining -= 64;
  }
  if (lane == 0)
    __hip_atomic_store(&status[t], status_word(tag, 2, prefix + local), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  return prefix;
}

// Global ids for the local nodes of a list of finished segments: an
// exclusive scan of (count[i] - sub) over i < n, offset by `base`, one
// kCountScanTile-entry tile per 1024-thread workgroup (8 consecutive entries
// per thread), joined by decoupled look-back on status[blockIdx.x] (with a
// tag no other use of `status` shares; workgroups only wait on lower ones,
// which the dispatcher starts first).  The last workgroup stores base +
// total to *total.  smem: 1024/64 + 2 words.
constexpr uint32_t kCountScanPer = 8;
constexpr uint32_t kCountScanTile = 1024 * kCountScanPer;
// A thread's kCountScanPer consecutive entries [i0, i0 + 8) of an array of
// n: two 16-byte words when all are in range and aligned, else one by one
// (absent entries read as 0, and are not stored).
__device__ __forceinline__ void load_per(const uint32_t *p, uint32_t i0, uint32_t n, uint32_t (&v)[kCountScanPer]) {
  static_assert(kCountScanPer == 8, "two 16-byte words");
  if (i0 + kCountScanPer <= n && ((uintptr_t)(p + i0) & 15u) == 0) {
    const uint4 x = reinterpret_cast<const uint4 *>(p + i0)[0];
    const uint4 y = reinterpret_cast<const uint4 *>(p + i0)[1];
    v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
    v[4] = y.x; v[5] = y.y; v[6] = y.z; v[7] = y.w;
  } else {
#pragma unroll
    for (uint32_t j = 0; j < kCountScanPer; j++) v[j] = i0 + j < n ? p[i0 + j] : 0u;
  }
}
__device__ __forceinline__ void store_per(uint32_t *p, uint32_t i0, uint32_t n, const uint32_t (&v)[kCountScanPer]) {
  if (i0 + kCountScanPer <= n && ((uintptr_t)(p + i0) & 15u) == 0) {
    reinterpret_cast<uint4 *>(p + i0)[0] = make_uint4(v[0], v[1], v[2], v[3]);
    reinterpret_cast<uint4 *>(p + i0)[1] = make_uint4(v[4], v[5], v[6], v[7]);
  } else {
#pragma unroll
    for (uint32_t j = 0; j < kCountScanPer; j++)
      if (i0 + j < n) p[i0 + j] = v[j];
  }
}
// One BLOCK-thread workgroup's part (tile blk of BLOCK*kCountScanPer
// entries) of such a scan; returns the total over all tiles to the last
// tile's thread 0 through *total when total is non-null.
template <int BLOCK>
__device__ inline void count_scan_part(const uint32_t *count, uint32_t *out, uint32_t n, uint32_t base,
                                       uint64_t *status, uint32_t tag, uint32_t blk, bool last, uint32_t *total,
                                       uint32_t *smem, uint32_t sub, uint32_t *err = nullptr) {
  const uint32_t i0 = blk * (BLOCK * kCountScanPer) + threadIdx.x * kCountScanPer;
  uint32_t v[kCountScanPer], sum = 0;
  load_per(count, i0, n, v);
#pragma unroll
  for (uint32_t j = 0; j < kCountScanPer; j++) {
    v[j] = i0 + j < n ? v[j] - sub : 0u;
    sum += v[j];
  }
  uint32_t tot;
  const uint32_t ex = block_excl_scan<BLOCK, true>(sum, smem, &tot);
  if (threadIdx.x < kWave) {
    const uint32_t prefix = lookback_exclusive_wave(status, blk, blk, tag, tot, false, err);
    if (threadIdx.x == 0) {
      smem[BLOCK / kWave + 1] = prefix;
      if (last && total) *total = base + prefix + tot;
    }
  }
  sync_block_lds();
  uint32_t o = base + smem[BLOCK / kWave + 1] + ex;
  uint32_t r[kCountScanPer];
#pragma unroll
  for (uint32_t j = 0; j < kCountScanPer; j++) {
    r[j] = o;
    o += v[j];
  }
  store_per(out, i0, n, r);
}
__device__ inline void count_scan_tile(const uint32_t *count, uint32_t *out, uint32_t n, uint32_t base,
                                       uint64_t *status, uint32_t tag, uint32_t *total, uint32_t *smem,
                                       uint32_t sub = 1, uint32_t *err = nullptr) {
  count_scan_part<1024>(count, out, n, base, status, tag, blockIdx.x, blockIdx.x == gridDim.x - 1, total, smem,
                        sub, err);
}

// block-wide min/max reduction of 12 floats (left lo/hi, right lo/hi)
template <int BLOCK, bool LDS_ONLY = false>
__device__ inline void block_reduce_bounds(float (&v)[12], float *smem /*12*BLOCK/64*/) {
  constexpr int NW = BLOCK / kWave;
  const uint32_t lane = lane_id(), wid = threadIdx.x / kWave;
#pragma unroll
  for (int j = 0; j < 12; j++) {
    bool isMin = (j % 6) < 3;
    v[j] = isMin ? wave_min(v[j]) : wave_max(v[j]);
  }
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < 12; j++) smem[j * NW + wid] = v[j];
  }
  block_sync<LDS_ONLY>();
  if (threadIdx.x < 12) {
    int j = threadIdx.x;
    bool isMin = (j % 6) < 3;
    float r = smem[j * NW];
    for (int w = 1; w < NW; w++) r = isMin ? fminf(r, smem[j * NW + w]) : fmaxf(r, smem[j * NW + w]);
    smem[j * NW] = r;
  }
  block_sync<LDS_ONLY>();
#pragma unroll
  for (int j = 0; j < 12; j++) v[j] = smem[j * NW];
  block_sync<LDS_ONLY>();
}

}  // namespace cubql
