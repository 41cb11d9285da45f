// block_ops.h -- wave64 / workgroup scan, rank and reduction primitives.
#pragma once
#include "common.h"

namespace cubql {

// inclusive wave scan (64 lanes)
__device__ inline uint32_t wave_incl_scan(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    uint32_t t = __shfl_up(v, o, kWave);
    if (lane >= (uint32_t)o) v += t;
  }
  return v;
}

template <bool LDS_ONLY>
__device__ __forceinline__ void block_sync() {
  if (LDS_ONLY) sync_block_lds();
  else __syncthreads();
}

// Exclusive scan across a workgroup of BLOCK threads.  `smem` needs
// BLOCK/64 + 1 words.  Returns the exclusive prefix; *total gets the sum.
// LDS_ONLY: barriers order LDS only (global stores stay in flight).
template <int BLOCK, bool LDS_ONLY = false>
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t *smem, uint32_t *total) {
  constexpr int NW = BLOCK / kWave;
  const uint32_t lane = lane_id(), wid = threadIdx.x / kWave;
  uint32_t incl = wave_incl_scan(v);
  if (lane == kWave - 1) smem[wid] = incl;
  block_sync<LDS_ONLY>();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int w = 0; w < NW; w++) {
      uint32_t c = smem[w];
      smem[w] = run;
      run += c;
    }
    smem[NW] = run;
  }
  block_sync<LDS_ONLY>();
  uint32_t r = smem[wid] + incl - v;
  *total = smem[NW];
  block_sync<LDS_ONLY>();
  return r;
}

// K exclusive block scans at once (one set of barriers): ex[q] = sum of
// v[q] over lower threads, tot[q] = the block's sum.  smem needs K*(NW+1)
// words.
template <int BLOCK, int K, bool LDS_ONLY = false>
__device__ inline void block_excl_scan_multi(const uint32_t (&v)[K], uint32_t (&ex)[K], uint32_t (&tot)[K],
                                             uint32_t *smem) {
  constexpr int NW = BLOCK / kWave;
  const uint32_t lane = lane_id(), wid = threadIdx.x / kWave;
  uint32_t incl[K];
#pragma unroll
  for (int q = 0; q < K; q++) {
    incl[q] = wave_incl_scan(v[q]);
    if (lane == kWave - 1) smem[q * (NW + 1) + wid] = incl[q];
  }
  block_sync<LDS_ONLY>();
  if (threadIdx.x < (uint32_t)K) {
    uint32_t *row = smem + threadIdx.x * (NW + 1);
    uint32_t run = 0;
    for (int w = 0; w < NW; w++) {
      const uint32_t c = row[w];
      row[w] = run;
      run += c;
    }
    row[NW] = run;
  }
  block_sync<LDS_ONLY>();
#pragma unroll
  for (int q = 0; q < K; q++) {
    ex[q] = smem[q * (NW + 1) + wid] + incl[q] - v[q];
    tot[q] = smem[q * (NW + 1) + NW];
  }
  block_sync<LDS_ONLY>();
}

// Ranks of flagged items in a striped tile: item k of thread t has tile
// index k*BLOCK + t.  rank[k] = number of flagged items with a smaller tile
// index.  smem needs ITEMS*BLOCK/64 words.  Returns the number of flagged
// items in the tile.
template <int BLOCK, int ITEMS, bool LDS_ONLY = false>
__device__ inline uint32_t block_rank_striped(const bool (&flag)[ITEMS],
                                              uint32_t (&rank)[ITEMS],
                                              uint32_t *smem) {
  constexpr int NW = BLOCK / kWave;
  constexpr int NC = ITEMS * NW;
  static_assert(NC <= kWave, "too many (item,wave) counters");
  const uint32_t lane = lane_id(), wid = threadIdx.x / kWave;
  uint64_t m[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; k++) {
    m[k] = __ballot(flag[k]);
    if (lane == 0) smem[k * NW + wid] = (uint32_t)__popcll(m[k]);
  }
  block_sync<LDS_ONLY>();
  if (wid == 0) {
    // each lane rewrites only the slot it read; smem[NC] is read after the
    // barrier below
    uint32_t c = lane < (uint32_t)NC ? smem[lane] : 0u;
    uint32_t incl = wave_incl_scan(c);
    if (lane < (uint32_t)NC) smem[lane] = incl - c;
    if (lane == (uint32_t)NC - 1) smem[NC] = incl;
  }
  block_sync<LDS_ONLY>();
#pragma unroll
  for (int k = 0; k < ITEMS; k++) rank[k] = smem[k * NW + wid] + mbcnt(m[k]);
  uint32_t tot = smem[NC];
  block_sync<LDS_ONLY>();
  return tot;
}

}  // namespace cubql
