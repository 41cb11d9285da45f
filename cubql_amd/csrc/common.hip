// common.hip -- error state, memory resource and small utility kernels.
#include <cstdarg>
#include <mutex>

#include "common.h"

namespace cubql {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
void clear_error() { g_err[0] = 0; }

// Default resource: stream-ordered allocation on the device's default pool
// with an unbounded release threshold, set once per device
// (AsyncGpuMemoryResource, builder/cuda.h:52-67).
static void init_default_pool() {
  static std::once_flag once[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return;
  std::call_once(once[dev], [dev]() {
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
      uint64_t thr = UINT64_MAX;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
  });
}

int mem_alloc(const cubql_mem_resource *mem, void **ptr, size_t bytes, hipStream_t s) {
  *ptr = nullptr;
  if (bytes == 0) bytes = 16;
  if (mem && mem->malloc_fn) {
    int st = mem->malloc_fn(ptr, bytes, (cubql_stream_t)s, mem->user);
    if (st != 0 || !*ptr) {
      set_error("memory resource failed to allocate %zu bytes (status %d)", bytes, st);
      return CUBQL_ERR_ALLOC;
    }
    return CUBQL_OK;
  }
  init_default_pool();
  hipError_t e = hipMallocAsync(ptr, bytes, s);
  if (e != hipSuccess || !*ptr) {
    set_error("hipMallocAsync(%zu) failed: %s", bytes, hipGetErrorString(e));
    return CUBQL_ERR_ALLOC;
  }
  return CUBQL_OK;
}

int mem_free(const cubql_mem_resource *mem, void *ptr, hipStream_t s) {
  if (!ptr) return CUBQL_OK;
  if (mem && mem->free_fn) {
    int st = mem->free_fn(ptr, (cubql_stream_t)s, mem->user);
    if (st != 0) {
      set_error("memory resource failed to free (status %d)", st);
      return CUBQL_ERR_ALLOC;
    }
    return CUBQL_OK;
  }
  hipError_t e = hipFreeAsync(ptr, s);
  if (e != hipSuccess) {
    set_error("hipFreeAsync failed: %s", hipGetErrorString(e));
    return CUBQL_ERR_HIP;
  }
  return CUBQL_OK;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

int fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s) {
  if (n == 0) return CUBQL_OK;
  uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, 4096);
  k_fill_u32<<<blocks, 256, 0, s>>>(p, v, n);
  CUBQL_LAUNCH_CHECK();
  return CUBQL_OK;
}

}  // namespace cubql

extern "C" const char *cubql_last_error_string(void) { return cubql::g_err; }
extern "C" const char *cubql_version(void) { return "cubql_amd 0.1 (gfx950)"; }

extern "C" int cubql_free_f3(cubql_bvh3f *bvh, cubql_stream_t s,
                             const cubql_mem_resource *mem) {
  using namespace cubql;
  clear_error();
  if (!bvh) {
    set_error("cubql_free_f3: bvh is NULL");
    return CUBQL_ERR_INVALID_ARG;
  }
  hipStream_t st = (hipStream_t)s;
  CUBQL_CHECK(mem_free(mem, bvh->primIDs, st));
  CUBQL_CHECK(mem_free(mem, bvh->nodes, st));
  CUBQL_HIP(hipStreamSynchronize(st));  // gpu_builder.h:259
  bvh->primIDs = nullptr;
  bvh->nodes = nullptr;
  bvh->numNodes = 0;
  bvh->numPrims = 0;
  return CUBQL_OK;
}
