// cuBQL/builder/cuda.h -- host builder API in the reference's spelling
// (builder/cuda.h:25-175) over the C ABI of libcubql_amd.so: the same
// names, argument order and defaults, with HIP streams in place of CUDA
// streams.  Errors: the reference prints and raises SIGINT on a device
// error (common.h:134-149) and throws std::runtime_error for unsupported
// cases (sah_builder.h:394); here every failure throws std::runtime_error
// carrying cubql_last_error_string().  Link libcubql_amd.so and
// libamdhip64.so.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "cuBQL/bvh.h"

namespace cuBQL {

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
inline cubql_bvh3f *c_bvh(bvh3f &b) { return &cuBQL::c_bvh(b); }
inline const cubql_box3f *c_box(const box3f *b) { return reinterpret_cast<const cubql_box3f *>(b); }
inline const cubql_build_config *c_cfg(const BuildConfig &c) {
  return reinterpret_cast<const cubql_build_config *>(&c);
}
}  // namespace detail

/*! gpuBuilder (builder/cuda.h:91-99; gpu_builder.h:31-67): spatial median
    (default) or, with buildConfig.enableSAH(), the 16-bin SAH builder. */
inline void gpuBuilder(bvh3f &bvh, const box3f *boxes, uint32_t numBoxes, BuildConfig buildConfig,
                       hipStream_t s = 0, GpuMemoryResource &memResource = defaultGpuMemResource()) {
  detail::Resource r(memResource);
  detail::check(cubql_gpu_builder_f3(detail::c_bvh(bvh), detail::c_box(boxes), numBoxes, detail::c_cfg(buildConfig), s, r.ptr),
                "cuBQL::gpuBuilder");
}

/*! gpuBuilder for a WideBVH<float,3,W>, W = 4 or 8 (builder/cuda.h:107-115;
    wide_gpu_builder.h:162-215): a binary build collapsed to width W. */
template <int W>
inline void gpuBuilder(WideBVH<float, 3, W> &bvh, const box3f *boxes, uint32_t numBoxes, BuildConfig buildConfig,
                       hipStream_t s = 0, GpuMemoryResource &memResource = defaultGpuMemResource()) {
  static_assert(W == 4 || W == 8, "WideBVH width must be 4 or 8");
  static_assert(sizeof(WideBVH<float, 3, W>) == sizeof(cubql_wbvh3f), "WideBVH layout");
  detail::Resource r(memResource);
  detail::check(cubql_gpu_builder_wide_f3(reinterpret_cast<cubql_wbvh3f *>(&bvh), W, detail::c_box(boxes), numBoxes,
                                          detail::c_cfg(buildConfig), s, r.ptr),
                "cuBQL::gpuBuilder(WideBVH)");
}

/*! refit (sm_builder.h:616-693): recompute every node box from prim boxes. */
inline void refit(bvh3f &bvh, const box3f *boxes, hipStream_t s = 0,
                  GpuMemoryResource &memResource = defaultGpuMemResource()) {
  detail::Resource r(memResource);
  detail::check(cubql_refit_f3(detail::c_bvh(bvh), detail::c_box(boxes), s, r.ptr), "cuBQL::refit");
}

namespace cuda {

/*! sahBuilder (builder/cuda.h:122-130); numBins extends the reference's
    fixed 16 (sah_builder.h:53-61). */
inline void sahBuilder(bvh3f &bvh, const box3f *boxes, uint32_t numPrims, BuildConfig buildConfig,
                       hipStream_t s = 0, GpuMemoryResource &memResource = defaultGpuMemResource(),
                       int numBins = 16) {
  detail::Resource r(memResource);
  detail::check(cubql_sah_builder_f3(detail::c_bvh(bvh), detail::c_box(boxes), numPrims, detail::c_cfg(buildConfig), numBins, s,
                                     r.ptr),
                "cuBQL::cuda::sahBuilder");
}

/*! radixBuilder (builder/cuda.h:134-140): 63-bit Morton keys. */
inline void radixBuilder(bvh3f &bvh, const box3f *boxes, uint32_t numPrims, BuildConfig buildConfig,
                         hipStream_t s = 0, GpuMemoryResource &memResource = defaultGpuMemResource()) {
  detail::Resource r(memResource);
  detail::check(cubql_radix_builder_f3(detail::c_bvh(bvh), detail::c_box(boxes), numPrims, detail::c_cfg(buildConfig), s, r.ptr),
                "cuBQL::cuda::radixBuilder");
}

/*! The 30-bit-key variant of radixBuilder (radix.h:34-38). */
inline void radix30Builder(bvh3f &bvh, const box3f *boxes, uint32_t numPrims, BuildConfig buildConfig,
                           hipStream_t s = 0, GpuMemoryResource &memResource = defaultGpuMemResource()) {
  detail::Resource r(memResource);
  detail::check(cubql_radix30_builder_f3(detail::c_bvh(bvh), detail::c_box(boxes), numPrims, detail::c_cfg(buildConfig), s, r.ptr),
                "cuBQL::cuda::radix30Builder");
}

/*! rebinRadixBuilder (builder/cuda.h:148-154). */
inline void rebinRadixBuilder(bvh3f &bvh, const box3f *boxes, uint32_t numPrims, BuildConfig buildConfig,
                              hipStream_t s = 0, GpuMemoryResource &memResource = defaultGpuMemResource()) {
  detail::Resource r(memResource);
  detail::check(cubql_rebin_radix_builder_f3(detail::c_bvh(bvh), detail::c_box(boxes), numPrims, detail::c_cfg(buildConfig), s,
                                             r.ptr),
                "cuBQL::cuda::rebinRadixBuilder");
}

/*! free (builder/cuda.h:155-161): release nodes and primIDs with the
    resource used to build. */
inline void free(bvh3f &bvh, hipStream_t s = 0, GpuMemoryResource &memResource = defaultGpuMemResource()) {
  detail::Resource r(memResource);
  detail::check(cubql_free_f3(detail::c_bvh(bvh), s, r.ptr), "cuBQL::cuda::free");
}

/*! free for a WideBVH (builder/cuda.h:164-170). */
template <int W>
inline void free(WideBVH<float, 3, W> &bvh, hipStream_t s = 0,
                 GpuMemoryResource &memResource = defaultGpuMemResource()) {
  detail::Resource r(memResource);
  detail::check(cubql_free_wide_f3(reinterpret_cast<cubql_wbvh3f *>(&bvh), s, r.ptr), "cuBQL::cuda::free(WideBVH)");
}

}  // namespace cuda

}  // namespace cuBQL
