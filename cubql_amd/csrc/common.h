// common.h -- internal helpers shared by the HIP sources of libcubql_amd.
//
// Layout types mirror the reference (bvh.h:71-111, math/box.h:60-196); the
// device helpers are wave64 (CDNA4) primitives.
#pragma once
#include <string.h>

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "cubql_amd/cubql_capi.h"

namespace cubql {

constexpr int kWave = 64;
constexpr uint64_t kOffsetMask = (1ull << 48) - 1;

struct __attribute__((aligned(16))) Node {
  float lo[3];
  float hi[3];
  uint64_t oc;  // offset:48 | count:16  (bvh.h:89-98)
};
static_assert(sizeof(Node) == 32, "Node must be 32 bytes");
static_assert(sizeof(cubql_node3f) == 32, "cubql_node3f must be 32 bytes");
static_assert(sizeof(cubql_bvh3f) == 32, "cubql_bvh3f must be 32 bytes");
static_assert(offsetof(cubql_bvh3f, primIDs) == 16, "bvh layout");
static_assert(sizeof(cubql_box3f) == 24, "box layout");
static_assert(sizeof(cubql_build_config) == 12, "BuildConfig layout");

__host__ __device__ inline uint64_t pack_oc(uint64_t off, uint64_t cnt) {
  return (off & kOffsetMask) | (cnt << 48);
}
__host__ __device__ inline uint64_t oc_offset(uint64_t oc) { return oc & kOffsetMask; }
__host__ __device__ inline uint32_t oc_count(uint64_t oc) { return (uint32_t)(oc >> 48); }

// The dispatcher hands workgroup b of a launch to XCD b % 8, and each of
// the 8 XCDs has its own L2.  For nb workgroups over work stored in a
// locality order (Morton-sorted queries), xcd_chunk(b, nb) is the index of
// the piece of work workgroup b takes such that each XCD gets one contiguous
// run of pieces (a bijection on [0, nb)), so an XCD's L2 serves one region
// of the tree instead of all of it.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_chunk(uint32_t b, uint32_t nb) {
  const uint32_t q = nb / kXcds, r = nb % kXcds, x = b % kXcds, j = b / kXcds;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + j;
}

// ---------------------------------------------------------------- errors
void set_error(const char *fmt, ...);
void clear_error();

#define CUBQL_HIP(call)                                                       \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      ::cubql::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,           \
                         hipGetErrorString(e_));                              \
      return CUBQL_ERR_HIP;                                                   \
    }                                                                         \
  } while (0)

#define CUBQL_CHECK(expr)                                                     \
  do {                                                                        \
    int st_ = (expr);                                                         \
    if (st_ != CUBQL_OK) return st_;                                          \
  } while (0)

#define CUBQL_LAUNCH_CHECK()                                                  \
  do {                                                                        \
    hipError_t e_ = hipGetLastError();                                        \
    if (e_ != hipSuccess) {                                                   \
      ::cubql::set_error("%s:%d kernel launch: %s", __FILE__, __LINE__,       \
                         hipGetErrorString(e_));                              \
      return CUBQL_ERR_HIP;                                                   \
    }                                                                         \
  } while (0)

// ---------------------------------------------------------- memory resource
// Allocations go through the caller's resource (GpuMemoryResource,
// builder/cuda.h:31-78); NULL means hipMallocAsync on the default pool.
int mem_alloc(const cubql_mem_resource *mem, void **ptr, size_t bytes, hipStream_t s);
int mem_free(const cubql_mem_resource *mem, void *ptr, hipStream_t s);

// RAII scratch list freed in reverse order on scope exit.
struct Scratch {
  const cubql_mem_resource *mem;
  hipStream_t s;
  void *ptrs[64];
  int n = 0;
  Scratch(const cubql_mem_resource *m, hipStream_t st) : mem(m), s(st) {}
  ~Scratch() {
    for (int i = n - 1; i >= 0; i--) mem_free(mem, ptrs[i], s);
  }
  template <typename T>
  int alloc(T **p, size_t count) {
    if (n >= 64) {
      set_error("scratch table full");
      return CUBQL_ERR_INTERNAL;
    }
    void *v = nullptr;
    int st = mem_alloc(mem, &v, count * sizeof(T) + 16, s);
    if (st != CUBQL_OK) return st;
    ptrs[n++] = v;
    *p = (T *)v;
    return CUBQL_OK;
  }
};

// Temporaries of one call carved from a single allocation, so the memory
// resource is called once instead of once per array: reserve() records each
// array, commit() allocates and assigns the pointers (256-byte aligned).
struct ScratchArena {
  const cubql_mem_resource *mem;
  hipStream_t s;
  void *slots[64];
  size_t offs[64];
  int n = 0;
  size_t total = 0;
  void *base = nullptr;
  ScratchArena(const cubql_mem_resource *m, hipStream_t st) : mem(m), s(st) {}
  ~ScratchArena() {
    if (base) mem_free(mem, base, s);
  }
  template <typename T>
  int reserve(T **p, size_t count) {
    if (n >= 64) {
      set_error("scratch table full");
      return CUBQL_ERR_INTERNAL;
    }
    slots[n] = (void *)p;
    offs[n] = total;
    n++;
    total += (count * sizeof(T) + 16 + 255) & ~(size_t)255;
    return CUBQL_OK;
  }
  int commit() {
    int st = mem_alloc(mem, &base, total, s);
    if (st != CUBQL_OK) return st;
    for (int i = 0; i < n; i++) {
      void *v = (char *)base + offs[i];
      memcpy(slots[i], &v, sizeof(void *));
    }
    return CUBQL_OK;
  }
};

// ----------------------------------------------------- float <-> ordered int
// AtomicBox encoding (builder_common.h:102-117): a monotone bijection from
// float to int32 so that min/max can use integer atomics.
__host__ __device__ inline int32_t enc_f(float f) {
  int32_t b = __float_as_int(f);
  return (b & (int32_t)0x80000000) ? (b ^ 0x7fffffff) : b;
}
__host__ __device__ inline float dec_f(int32_t b) {
  return __int_as_float((b & (int32_t)0x80000000) ? (b ^ 0x7fffffff) : b);
}

// ------------------------------------------------------------ wave helpers
__device__ inline float wave_min(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, kWave));
  return v;
}
__device__ inline float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, kWave));
  return v;
}
__device__ inline uint32_t lane_id() { return __lane_id(); }
// number of set bits of m below this lane
__device__ inline uint32_t mbcnt(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Workgroup barrier that orders LDS only.  __syncthreads() also waits for
// every outstanding global access of the calling wave (s_waitcnt vmcnt(0)
// before the barrier), which stalls loops that stream stores; use this where
// threads communicate through LDS alone.
__device__ __forceinline__ void sync_block_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
// The same for a one-wave workgroup: a wave's LDS accesses execute in
// program order, so only the compiler must not reorder them.
__device__ __forceinline__ void sync_wave_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__host__ __device__ inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// device-side fill (used instead of hipMemsetAsync so the sequence stays
// capturable in a graph)
int fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s);

}  // namespace cubql
