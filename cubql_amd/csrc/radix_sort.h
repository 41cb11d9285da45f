// radix_sort.h -- stable LSD key/value radix sort for gfx950 (onesweep).
//
// Replaces the cub::DeviceRadixSort::SortPairs calls of the reference
// builders (builder/cuda/radix.h:666-687, rebinMortonBuilder.h:1242-1263):
// keys (uint32 or uint64) with uint32 values, ascending, stable.
//
// Design: 8-bit digits.  One histogram pass computes the digit counts of
// every sort pass at once (per-workgroup histograms, summed in chunks with
// one atomic per digit and chunk).  Each sort pass is then a single read + single write of the
// data: a workgroup claims a 4096-key tile (ticket order, so it only waits
// on tiles that running workgroups hold), ranks keys per digit with wave64
// ballots, resolves each digit's global offset by a decoupled look-back over
// earlier tiles, stages the tile in LDS in digit order and writes it out so
// that consecutive lanes write consecutive addresses.
#pragma once
#include "common.h"

namespace cubql {

// Sorts (kA, vA) of *nDev (device-side count, <= nMax) pairs on key bits
// [0, keyBits).  (kB, vB) is the ping-pong buffer; *kRes / *vRes receive the
// buffers holding the result (kA/vA when the pass count is even).  err: a
// device word the caller checks after the sort; when given, the look-back
// waits are bounded and a wait past the bound sets it (the result is then
// invalid); when NULL they are not bounded.
int radix_sort_pairs_dev_u32(uint32_t *kA, uint32_t *vA, uint32_t *kB, uint32_t *vB,
                             const uint32_t *nDev, uint32_t nMax, int keyBits, hipStream_t s,
                             const cubql_mem_resource *mem, uint32_t **kRes, uint32_t **vRes,
                             uint32_t *err = nullptr);
int radix_sort_pairs_dev_u64(uint64_t *kA, uint32_t *vA, uint64_t *kB, uint32_t *vB,
                             const uint32_t *nDev, uint32_t nMax, int keyBits, hipStream_t s,
                             const cubql_mem_resource *mem, uint64_t **kRes, uint32_t **vRes,
                             uint32_t *err = nullptr);

// Workspace-based variant (no allocation; capturable in a graph).
size_t radix_sort_ws_bytes(uint32_t nMax, int keyBits);
int radix_sort_pairs_ws_u32(uint32_t *kA, uint32_t *vA, uint32_t *kB, uint32_t *vB,
                            const uint32_t *nDev, uint32_t nMax, int keyBits, hipStream_t s,
                            void *ws, uint32_t **kRes, uint32_t **vRes);
int radix_sort_pairs_ws_u64(uint64_t *kA, uint32_t *vA, uint64_t *kB, uint32_t *vB,
                            const uint32_t *nDev, uint32_t nMax, int keyBits, hipStream_t s,
                            void *ws, uint64_t **kRes, uint32_t **vRes);

}  // namespace cubql
