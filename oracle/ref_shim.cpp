// oracle/ref_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A thin extern "C" wrapper around the reference's own host path, compiled
// from the reference headers where they lie (see oracle/Makefile; nothing of
// the reference is copied here).  Produces oracle/_ref/libcubql_ref.so, used
// to pin the oracle restatement and as the CPU baseline.  It exposes:
//   - cuBQL::cpu::spatialMedian            (builder/cpu/spatialMedian.h)
//   - cuBQL::points::findClosest / findKNN (queries/points/*.h) on the host
//   - samples::UniformPointGenerator       (samples/common/Generator.cpp)
#include "cuBQL/bvh.h"
#include "cuBQL/builder/cpu.h"
#include "cuBQL/queries/points/findClosest.h"
#include "cuBQL/queries/points/knn.h"
#include "samples/common/Generator.h"

#include <chrono>
#include <cstring>
#include <vector>

using namespace cuBQL;

static_assert(sizeof(bvh3f::Node) == 32, "node layout");

extern "C" {

int ref_uniform_points(int seed, int n, double lo, double hi, float *out) {
  vec_t<double, 3> vlo(lo), vhi(hi);
  samples::UniformPointGenerator<3> gen{vlo, vhi};
  std::vector<vec3d> pts = gen.generate(n, seed);
  std::vector<vec3f> f = samples::convert<float>(pts);
  std::memcpy(out, f.data(), sizeof(vec3f) * (size_t)n);
  return 0;
}

// Builds with cpu::spatialMedian; copies nodes/primIDs out (caller provides
// capacity 2n+2 nodes and n primIDs).
int ref_cpu_build(const float *boxes, int n, int leafThreshold, void *nodesOut,
                  uint32_t *primIDsOut, int *numNodes, int *numPrims) {
  bvh3f bvh;
  BuildConfig cfg;
  cfg.makeLeafThreshold = leafThreshold;
  cpu::spatialMedian(bvh, (const box3f *)boxes, (uint32_t)n, cfg);
  std::memcpy(nodesOut, bvh.nodes, sizeof(bvh3f::Node) * bvh.numNodes);
  std::memcpy(primIDsOut, bvh.primIDs, sizeof(uint32_t) * bvh.numPrims);
  *numNodes = (int)bvh.numNodes;
  *numPrims = (int)bvh.numPrims;
  cpu::freeBVH(bvh);
  return 0;
}

static bvh3f make_view(const void *nodes, int numNodes, const uint32_t *primIDs, int numPrims) {
  bvh3f bvh;
  bvh.nodes = (bvh3f::Node *)nodes;
  bvh.numNodes = (uint32_t)numNodes;
  bvh.primIDs = (uint32_t *)primIDs;
  bvh.numPrims = (uint32_t)numPrims;
  return bvh;
}

void ref_fcp(const void *nodes, int numNodes, const uint32_t *primIDs, int numPrims,
             const float *points, const float *queries, int nq, float sqrMax,
             int32_t *ids, float *sqrDist) {
  bvh3f bvh = make_view(nodes, numNodes, primIDs, numPrims);
  const vec3f *P = (const vec3f *)points;
  const vec3f *Q = (const vec3f *)queries;
  for (int i = 0; i < nq; i++) {
    int id = points::findClosest(bvh, P, Q[i], sqrMax);
    ids[i] = id;
    sqrDist[i] = id >= 0 ? fSqrDistance_rd(P[id], Q[i]) : INFINITY;
  }
}

void ref_knn(const void *nodes, int numNodes, const uint32_t *primIDs, int numPrims,
             const float *points, const float *queries, int nq, int k, float sqrMax,
             int32_t *ids, float *sqrDist, int32_t *numFound) {
  bvh3f bvh = make_view(nodes, numNodes, primIDs, numPrims);
  const vec3f *P = (const vec3f *)points;
  const vec3f *Q = (const vec3f *)queries;
  std::vector<knn::Candidate> cand(k);
  for (int i = 0; i < nq; i++) {
    knn::Result r = points::findKNN(cand.data(), k, bvh, P, Q[i], sqrMax);
    for (int j = 0; j < k; j++) {
      ids[(size_t)i * k + j] = j < r.numFound ? cand[j].primID : -1;
      sqrDist[(size_t)i * k + j] = j < r.numFound ? cand[j].sqrDist : INFINITY;
    }
    numFound[i] = r.numFound;
  }
}

// Reference-style timing of the serial CPU builder (buildBench.h:115-133
// method is applied by the caller); returns seconds for one build.
double ref_time_cpu_build(const float *boxes, int n, int leafThreshold) {
  auto t0 = std::chrono::steady_clock::now();
  bvh3f bvh;
  BuildConfig cfg;
  cfg.makeLeafThreshold = leafThreshold;
  cpu::spatialMedian(bvh, (const box3f *)boxes, (uint32_t)n, cfg);
  auto t1 = std::chrono::steady_clock::now();
  cpu::freeBVH(bvh);
  return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
