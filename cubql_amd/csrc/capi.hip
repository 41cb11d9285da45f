// capi.hip -- builder entry points of the C ABI (include/cubql_amd/cubql_capi.h).
#include "common.h"
#include "refit.h"

using namespace cubql;

static int check_build_args(const char *fn, cubql_bvh3f *bvh, const cubql_box3f *boxes, uint32_t n,
                            const cubql_build_config *cfg) {
  if (!bvh || !cfg || (n && !boxes)) {
    set_error("%s: null argument", fn);
    return CUBQL_ERR_INVALID_ARG;
  }
  if (n > 0x7fffffffu) {
    set_error("%s: %u prims exceeds 2^31-1", fn, n);
    return CUBQL_ERR_INVALID_ARG;
  }
  return CUBQL_OK;
}

// makeLeafThreshold 0 -> 1 (gpu_builder.h:60-63); the 16-bit node count field
// (bvh.h:79) caps leaves at 65535 prims.
static uint32_t leaf_threshold(const cubql_build_config *cfg) {
  int t = cfg->makeLeafThreshold > 0 ? cfg->makeLeafThreshold : 1;
  return (uint32_t)std::min(t, 65535);
}

// BuildConfig::maxAllowedLeafSize (bvh.h:50-53); <= 0 means only the 16-bit
// count limit applies.
static uint32_t leaf_cap(const cubql_build_config *cfg) {
  return cfg->maxAllowedLeafSize > 0 ? (uint32_t)std::min(cfg->maxAllowedLeafSize, 65535) : 65535u;
}

extern "C" int cubql_gpu_builder_f3(cubql_bvh3f *bvh, const cubql_box3f *d_boxes, uint32_t numBoxes,
                                    const cubql_build_config *cfg, cubql_stream_t stream,
                                    const cubql_mem_resource *mem) {
  clear_error();
  CUBQL_CHECK(check_build_args("cubql_gpu_builder_f3", bvh, d_boxes, numBoxes, cfg));
  if (numBoxes == 0) return CUBQL_OK;  // gpu_builder.h:39
  hipStream_t s = (hipStream_t)stream;
  switch (cfg->buildMethod) {
    case CUBQL_SPATIAL_MEDIAN:
      return sm_build(bvh, d_boxes, numBoxes, leaf_threshold(cfg), s, mem);
    case CUBQL_SAH:
      return sah_build(bvh, d_boxes, numBoxes, leaf_threshold(cfg), leaf_cap(cfg), 16, s, mem);
    case CUBQL_ELH:
      // elhBuilder (elh_builder.h:437-557): 8 bins, edge-length cost; a
      // makeLeafThreshold of 0 and of 1 give the same tree there (a 1-prim
      // node has no split with two non-empty sides)
      return sah_build(bvh, d_boxes, numBoxes, leaf_threshold(cfg), leaf_cap(cfg), 8, s, mem, true);
    default:
      set_error("cubql_gpu_builder_f3: buildMethod %d not supported", cfg->buildMethod);
      return CUBQL_ERR_UNSUPPORTED;
  }
}

extern "C" int cubql_sah_builder_f3(cubql_bvh3f *bvh, const cubql_box3f *d_boxes, uint32_t numPrims,
                                    const cubql_build_config *cfg, int numBins, cubql_stream_t stream,
                                    const cubql_mem_resource *mem) {
  clear_error();
  CUBQL_CHECK(check_build_args("cubql_sah_builder_f3", bvh, d_boxes, numPrims, cfg));
  if (numPrims == 0) return CUBQL_OK;
  return sah_build(bvh, d_boxes, numPrims, leaf_threshold(cfg), leaf_cap(cfg), numBins,
                   (hipStream_t)stream, mem);
}
