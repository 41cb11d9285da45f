// radix_sort.h -- stable LSD key/value radix sort for gfx950 (onesweep).
//
// Replaces the cub::DeviceRadixSort::SortPairs calls of the reference
// builders (builder/cuda/radix.h:666-687, rebinMortonBuilder.h:1242-1263):
// keys (uint32 or uint64) with uint32 values, ascending, stable.
//
// Design: 8-bit digits.  One histogram pass computes the digit counts of
// every sort pass at once (per-workgroup histograms, reduced without global
// atomics).  Each sort pass is then a single read + single write of the
// data: a workgroup claims a 4096-key tile (ticket order), ranks keys per
// digit with wave64 ballots, resolves each digit's global offset by a
// decoupled look-back over earlier tiles, stages the tile in LDS in digit
// order and writes it out so that consecutive lanes write consecutive
// addresses.
#pragma once
#include "common.h"

namespace cubql {

// Sort (keys, vals) of length n.  Scratch comes from `mem`.  The sorted
// result is written to (keysOut, valsOut); keysIn/valsIn are clobbered.
// Only bits [0, keyBits) are sorted on.
int radix_sort_pairs_u32(uint32_t *keysIn, uint32_t *valsIn, uint32_t *keysOut, uint32_t *valsOut,
                         uint32_t n, int keyBits, hipStream_t s, const cubql_mem_resource *mem);
int radix_sort_pairs_u64(uint64_t *keysIn, uint32_t *valsIn, uint64_t *keysOut, uint32_t *valsOut,
                         uint32_t n, int keyBits, hipStream_t s, const cubql_mem_resource *mem);

}  // namespace cubql
