// cubql_cpp_check.cpp -- exercises the C++ host API (include/cubql_amd/cubql.hpp)
// the way a reference user would call cuBQL::gpuBuilder & co.
//
//   cubql_cpp_check abi   no GPU work: version string, argument errors
//                         surface as std::runtime_error
//   cubql_cpp_check gpu   builds with every builder (default and a counting
//                         user resource), runs closest-point queries and
//                         compares them with a float brute force (lowest id
//                         on ties); checks that free() returns every byte
// Exit status 0 on success.
#include "cubql_amd/cubql.hpp"

#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

namespace {

struct CountingResource final : cuBQL::GpuMemoryResource {
  std::map<void *, size_t> live;
  size_t allocs = 0;
  hipError_t malloc(void **ptr, size_t size, hipStream_t) override {
    hipError_t e = hipMalloc(ptr, size);
    if (e == hipSuccess) { live[*ptr] = size; allocs++; }
    return e;
  }
  hipError_t free(void *ptr, hipStream_t s) override {
    if (hipStreamSynchronize(s) != hipSuccess) return hipErrorUnknown;
    live.erase(ptr);
    return hipFree(ptr);
  }
};

// small deterministic generator (not the reference's; any fixed input works)
std::vector<float> points(size_t n, uint32_t seed) {
  std::vector<float> p(3 * n);
  uint32_t s = seed * 747796405u + 2891336453u;
  for (auto &v : p) {
    s = s * 1664525u + 1013904223u;
    v = (float)(s >> 8) / (float)(1u << 24);
  }
  return p;
}

int fail(const char *what) {
  std::fprintf(stderr, "FAIL: %s\n", what);
  return 1;
}

#define HIPOK(x) do { if ((x) != hipSuccess) return fail(#x); } while (0)

int run_abi() {
  std::printf("version %s\n", cubql_version());
  cuBQL::bvh3f bvh;
  bool threw = false;
  try {
    cuBQL::gpuBuilder(bvh, nullptr, 5, cuBQL::BuildConfig());
  } catch (const std::runtime_error &e) {
    threw = std::strstr(e.what(), "cuBQL::gpuBuilder") != nullptr;
    std::printf("expected error: %s\n", e.what());
  }
  if (!threw) return fail("null boxes did not throw");
  std::printf("OK abi\n");
  return 0;
}

int run_gpu() {
  const size_t n = 20000, nq = 2000;
  std::vector<float> P = points(n, 1), Q = points(nq, 2);
  std::vector<cuBQL::box3f> boxes(n);
  for (size_t i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) boxes[i].lower[d] = boxes[i].upper[d] = P[3 * i + d];
  cuBQL::box3f *dBoxes;
  float *dP, *dQ, *dD;
  int32_t *dI;
  HIPOK(hipMalloc(&dBoxes, n * sizeof(cuBQL::box3f)));
  HIPOK(hipMalloc(&dP, n * 12));
  HIPOK(hipMalloc(&dQ, nq * 12));
  HIPOK(hipMalloc(&dI, nq * 4));
  HIPOK(hipMalloc(&dD, nq * 4));
  HIPOK(hipMemcpy(dBoxes, boxes.data(), n * sizeof(cuBQL::box3f), hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(dP, P.data(), n * 12, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(dQ, Q.data(), nq * 12, hipMemcpyHostToDevice));
  // brute force, (dx*dx + dy*dy) + dz*dz without contraction, lowest id
  std::vector<int32_t> refI(nq);
  std::vector<float> refD(nq);
  for (size_t q = 0; q < nq; q++) {
    float best = INFINITY;
    int32_t bi = -1;
    for (size_t i = 0; i < n; i++) {
      const float dx = P[3 * i] - Q[3 * q], dy = P[3 * i + 1] - Q[3 * q + 1], dz = P[3 * i + 2] - Q[3 * q + 2];
      const float dd = dx * dx + dy * dy + dz * dz;
      if (dd < best) { best = dd; bi = (int32_t)i; }
    }
    refI[q] = bi;
    refD[q] = best;
  }
  CountingResource counting;
  cuBQL::BuildConfig cfg;
  cfg.makeLeafThreshold = 8;
  struct Case { const char *name; int which; bool user; };
  const Case cases[] = {{"gpuBuilder", 0, false}, {"gpuBuilder(SAH)", 1, true}, {"radixBuilder", 2, true},
                        {"radix30Builder", 3, false}, {"rebinRadixBuilder", 4, true}, {"sahBuilder(32)", 5, false}};
  for (const Case &cs : cases) {
    cuBQL::GpuMemoryResource &mem = cs.user ? (cuBQL::GpuMemoryResource &)counting : cuBQL::defaultGpuMemResource();
    cuBQL::bvh3f bvh;
    cuBQL::BuildConfig c = cfg;
    switch (cs.which) {
      case 0: cuBQL::gpuBuilder(bvh, dBoxes, n, c, 0, mem); break;
      case 1: cuBQL::gpuBuilder(bvh, dBoxes, n, c.enableSAH(), 0, mem); break;
      case 2: cuBQL::cuda::radixBuilder(bvh, dBoxes, n, c, 0, mem); break;
      case 3: cuBQL::cuda::radix30Builder(bvh, dBoxes, n, c, 0, mem); break;
      case 4: cuBQL::cuda::rebinRadixBuilder(bvh, dBoxes, n, c, 0, mem); break;
      default: cuBQL::cuda::sahBuilder(bvh, dBoxes, n, c, 0, mem, 32); break;
    }
    cuBQL::points::findClosest(bvh, dP, dQ, nq, dI, dD);
    std::vector<int32_t> gi(nq);
    std::vector<float> gd(nq);
    HIPOK(hipMemcpy(gi.data(), dI, nq * 4, hipMemcpyDeviceToHost));
    HIPOK(hipMemcpy(gd.data(), dD, nq * 4, hipMemcpyDeviceToHost));
    for (size_t q = 0; q < nq; q++)
      if (gi[q] != refI[q] || gd[q] != refD[q]) return fail(cs.name);
    cuBQL::cuda::free(bvh, 0, mem);
    std::printf("OK %s: %u nodes\n", cs.name, bvh.numNodes);
  }
  HIPOK(hipDeviceSynchronize());
  if (counting.allocs == 0) return fail("user resource never used");
  if (!counting.live.empty()) return fail("user resource: memory not returned");
  std::printf("OK user resource: %zu allocations, all freed\n", counting.allocs);
  HIPOK(hipFree(dBoxes));
  HIPOK(hipFree(dP));
  HIPOK(hipFree(dQ));
  HIPOK(hipFree(dI));
  HIPOK(hipFree(dD));
  std::printf("OK gpu\n");
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  const char *mode = argc > 1 ? argv[1] : "abi";
  try {
    if (!std::strcmp(mode, "abi")) return run_abi();
    if (!std::strcmp(mode, "gpu")) return run_gpu();
  } catch (const std::exception &e) {
    std::fprintf(stderr, "FAIL: exception %s\n", e.what());
    return 1;
  }
  return fail("unknown mode");
}
