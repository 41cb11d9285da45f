// refit.h -- internal entry of the generic bottom-up refit (refit.hip).
#pragma once
#include "common.h"

namespace cubql {
// Recomputes every reachable node box of *bvh from the prim boxes.
int refit_tree(cubql_bvh3f *bvh, const cubql_box3f *boxes, hipStream_t s,
               const cubql_mem_resource *mem);
// Spatial-median builder (sm_builder.hip); T = leaf threshold (>= 1).
int sm_build(cubql_bvh3f *bvh, const cubql_box3f *boxes, uint32_t n, uint32_t T, hipStream_t s,
             const cubql_mem_resource *mem);
// Binned SAH builder (sah_builder.hip); 1 <= T <= cap <= 65535 after
// clamping, numBins in [2, 64].  elh: the edge-length heuristic's bin rule
// and box measure instead of SAH's (elh_builder.h).
int sah_build(cubql_bvh3f *bvh, const cubql_box3f *boxes, uint32_t n, uint32_t T, uint32_t cap,
              int numBins, hipStream_t s, const cubql_mem_resource *mem, bool elh = false);
}  // namespace cubql
