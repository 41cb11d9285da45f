// cubql.hpp -- host C++ API in the reference's spelling (namespace cuBQL)
// over the C ABI of libcubql_amd.so (cubql_capi.h).
//
// Mirrors builder/cuda.h:25-175 and bvh.h:25-111 for BinaryBVH<float,3>:
// the same types, names, argument order and defaults, with HIP streams in
// place of CUDA streams.  Errors: the reference prints and raises SIGINT on
// any device error (common.h:134-149) and throws std::runtime_error for
// unsupported cases (sah_builder.h:394); here every failure throws
// std::runtime_error carrying cubql_last_error_string().
//
// Host-only header: include from code compiled by hipcc or any C++17
// compiler with the ROCm include path and __HIP_PLATFORM_AMD__; link
// libcubql_amd.so and libamdhip64.so.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "cubql_amd/cubql_capi.h"

namespace cuBQL {

using box3f = cubql_box3f;

/*! BuildConfig (bvh.h:31-64), byte-identical to cubql_build_config. */
struct BuildConfig {
  enum BuildMethod { SPATIAL_MEDIAN = 0, SAH, ELH };
  BuildConfig &enableSAH() { buildMethod = SAH; return *this; }
  BuildConfig &enableELH() { buildMethod = ELH; return *this; }
  int maxAllowedLeafSize = 1 << 15;
  int makeLeafThreshold = 0;
  BuildMethod buildMethod = SPATIAL_MEDIAN;
};
static_assert(sizeof(BuildConfig) == sizeof(cubql_build_config), "BuildConfig layout");

template <typename T, int D> struct BinaryBVH;

/*! BinaryBVH<float,3> (bvh.h:71-111), byte-identical to cubql_bvh3f: node 0
    is the root, a node with count 0 has its children at offset and
    offset+1, a leaf holds primIDs[offset .. offset+count). */
template <> struct BinaryBVH<float, 3> {
  using Node = cubql_node3f;
  Node *nodes = nullptr;
  uint32_t numNodes = 0;
  uint32_t *primIDs = nullptr;
  uint32_t numPrims = 0;
};
static_assert(sizeof(BinaryBVH<float, 3>) == sizeof(cubql_bvh3f), "BinaryBVH layout");
using bvh3f = BinaryBVH<float, 3>;

/*! GpuMemoryResource (builder/cuda.h:31-34): every allocation a build makes
    (outputs and scratch) goes through it; free the BVH with the same one. */
struct GpuMemoryResource {
  virtual ~GpuMemoryResource() = default;
  virtual hipError_t malloc(void **ptr, size_t size, hipStream_t s) = 0;
  virtual hipError_t free(void *ptr, hipStream_t s) = 0;
};

/*! The default resource (builder/cuda.h:52-78): stream-ordered allocation
    on the device's default pool, whose release threshold the library sets to
    unbounded.  Passing it selects the library's built-in allocator. */
struct AsyncGpuMemoryResource final : GpuMemoryResource {
  hipError_t malloc(void **ptr, size_t size, hipStream_t s) override { return hipMallocAsync(ptr, size, s); }
  hipError_t free(void *ptr, hipStream_t s) override { return hipFreeAsync(ptr, s); }
};

inline GpuMemoryResource &defaultGpuMemResource() {
  static AsyncGpuMemoryResource r;
  return r;
}

namespace detail {
inline int malloc_thunk(void **p, size_t n, cubql_stream_t s, void *user) {
  return static_cast<GpuMemoryResource *>(user)->malloc(p, n, (hipStream_t)s) == hipSuccess ? 0 : -1;
}
inline int free_thunk(void *p, cubql_stream_t s, void *user) {
  return static_cast<GpuMemoryResource *>(user)->free(p, (hipStream_t)s) == hipSuccess ? 0 : -1;
}
// NULL selects the library's own default pool
struct Resource {
  cubql_mem_resource table{};
  const cubql_mem_resource *ptr = nullptr;
  explicit Resource(GpuMemoryResource &m) {
    if (&m != &defaultGpuMemResource()) {
      table.malloc_fn = &malloc_thunk;
      table.free_fn = &free_thunk;
      table.user = &m;
      ptr = &table;
    }
  }
};
inline void check(int status, const char *fn) {
  if (status != CUBQL_OK) throw std::runtime_error(std::string(fn) + ": " + cubql_last_error_string());
}
inline cubql_bvh3f *c_bvh(bvh3f &b) { return reinterpret_cast<cubql_bvh3f *>(&b); }
inline const cubql_build_config *c_cfg(const BuildConfig &c) {
  return reinterpret_cast<const cubql_build_config *>(&c);
}
}  // namespace detail

/*! gpuBuilder (builder/cuda.h:91-99; gpu_builder.h:31-67): spatial median
    (default) or, with buildConfig.enableSAH(), the 16-bin SAH builder. */
inline void gpuBuilder(bvh3f &bvh, const box3f *boxes, uint32_t numBoxes, BuildConfig buildConfig,
                       hipStream_t s = 0, GpuMemoryResource &memResource = defaultGpuMemResource()) {
  detail::Resource r(memResource);
  detail::check(cubql_gpu_builder_f3(detail::c_bvh(bvh), boxes, numBoxes, detail::c_cfg(buildConfig), s, r.ptr),
                "cuBQL::gpuBuilder");
}

/*! refit (sm_builder.h:616-693): recompute every node box from prim boxes. */
inline void refit(bvh3f &bvh, const box3f *boxes, hipStream_t s = 0,
                  GpuMemoryResource &memResource = defaultGpuMemResource()) {
  detail::Resource r(memResource);
  detail::check(cubql_refit_f3(detail::c_bvh(bvh), boxes, s, r.ptr), "cuBQL::refit");
}

namespace cuda {

/*! sahBuilder (builder/cuda.h:122-130); numBins extends the reference's
    fixed 16 (sah_builder.h:53-61). */
inline void sahBuilder(bvh3f &bvh, const box3f *boxes, uint32_t numPrims, BuildConfig buildConfig,
                       hipStream_t s = 0, GpuMemoryResource &memResource = defaultGpuMemResource(),
                       int numBins = 16) {
  detail::Resource r(memResource);
  detail::check(cubql_sah_builder_f3(detail::c_bvh(bvh), boxes, numPrims, detail::c_cfg(buildConfig), numBins, s,
                                     r.ptr),
                "cuBQL::cuda::sahBuilder");
}

/*! radixBuilder (builder/cuda.h:134-140): 63-bit Morton keys. */
inline void radixBuilder(bvh3f &bvh, const box3f *boxes, uint32_t numPrims, BuildConfig buildConfig,
                         hipStream_t s = 0, GpuMemoryResource &memResource = defaultGpuMemResource()) {
  detail::Resource r(memResource);
  detail::check(cubql_radix_builder_f3(detail::c_bvh(bvh), boxes, numPrims, detail::c_cfg(buildConfig), s, r.ptr),
                "cuBQL::cuda::radixBuilder");
}

/*! The 30-bit-key variant of radixBuilder (radix.h:34-38). */
inline void radix30Builder(bvh3f &bvh, const box3f *boxes, uint32_t numPrims, BuildConfig buildConfig,
                           hipStream_t s = 0, GpuMemoryResource &memResource = defaultGpuMemResource()) {
  detail::Resource r(memResource);
  detail::check(cubql_radix30_builder_f3(detail::c_bvh(bvh), boxes, numPrims, detail::c_cfg(buildConfig), s, r.ptr),
                "cuBQL::cuda::radix30Builder");
}

/*! rebinRadixBuilder (builder/cuda.h:148-154). */
inline void rebinRadixBuilder(bvh3f &bvh, const box3f *boxes, uint32_t numPrims, BuildConfig buildConfig,
                              hipStream_t s = 0, GpuMemoryResource &memResource = defaultGpuMemResource()) {
  detail::Resource r(memResource);
  detail::check(cubql_rebin_radix_builder_f3(detail::c_bvh(bvh), boxes, numPrims, detail::c_cfg(buildConfig), s,
                                             r.ptr),
                "cuBQL::cuda::rebinRadixBuilder");
}

/*! free (builder/cuda.h:155-161): release nodes and primIDs with the
    resource used to build. */
inline void free(bvh3f &bvh, hipStream_t s = 0, GpuMemoryResource &memResource = defaultGpuMemResource()) {
  detail::Resource r(memResource);
  detail::check(cubql_free_f3(detail::c_bvh(bvh), s, r.ptr), "cuBQL::cuda::free");
}

}  // namespace cuda

/*! Batched host-callable forms of the device queries: one query per
    element, results in device arrays (see cubql_capi.h for semantics). */
namespace points {
/*! points::findClosest (queries/points/findClosest.h:49-62): id of the
    closest point within sqrt(sqrMaxSearchRadius) (-1 if none), ties to the
    lowest primID; sqrDist may be null. */
inline void findClosest(const bvh3f &bvh, const float *d_points, const float *d_queries, uint32_t numQueries,
                        int32_t *d_ids, float *d_sqrDist = nullptr, float sqrMaxSearchRadius = INFINITY,
                        hipStream_t s = 0) {
  detail::check(cubql_fcp_points_f3(reinterpret_cast<const cubql_bvh3f *>(&bvh), d_points, d_queries, numQueries,
                                    sqrMaxSearchRadius, d_ids, d_sqrDist, s),
                "cuBQL::points::findClosest");
}
/*! points::findKNN (queries/points/knn.h:49-68): k <= 64 nearest per query,
    ordered by (distance, primID); unfilled slots -1 / +inf. */
inline void findKNN(const bvh3f &bvh, const float *d_points, const float *d_queries, uint32_t numQueries, int k,
                    int32_t *d_ids, float *d_sqrDist, int32_t *d_numFound = nullptr,
                    float sqrMaxSearchRadius = INFINITY, hipStream_t s = 0) {
  detail::check(cubql_knn_points_f3(reinterpret_cast<const cubql_bvh3f *>(&bvh), d_points, d_queries, numQueries, k,
                                    sqrMaxSearchRadius, d_ids, d_sqrDist, d_numFound, s),
                "cuBQL::points::findKNN");
}
}  // namespace points

namespace triangles {
/*! triangles::findClosest (queries/triangles/findClosest.h:132-152) for a
    mesh (d_tris = int3[] or null for the non-indexed 3-vertex layout). */
inline void findClosest(const bvh3f &bvh, const int32_t *d_tris, const float *d_verts, const float *d_queries,
                        uint32_t numQueries, int32_t *d_triIDs, float *d_sqrDist, float *d_closest = nullptr,
                        float sqrMaxSearchRadius = INFINITY, hipStream_t s = 0) {
  detail::check(cubql_fcp_triangles_f3(reinterpret_cast<const cubql_bvh3f *>(&bvh), d_tris, d_verts, d_queries,
                                       numQueries, sqrMaxSearchRadius, d_triIDs, d_sqrDist, d_closest, s),
                "cuBQL::triangles::findClosest");
}
}  // namespace triangles

}  // namespace cuBQL
