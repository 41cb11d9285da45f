// device_api_sample.hip -- user kernels written against the reference-
// spelled device headers (include/cuBQL/...), shaped like the reference's
// samples/s01_closestPoint_points_gpu/closestPoint.cu:31-66 and
// samples/s02_distanceToTriangleMesh/distanceToTriangleMesh.cu, with HIP in
// place of CUDA.
//
//   device_api_sample <dir>
// reads <dir>/{data,queries,verts}.f32 (n x 3 float32) and <dir>/tris.i32
// (t x 3 int32), builds with cuBQL::gpuBuilder, runs the kernels and writes
// <dir>/{fcp.i32, knn_ids.i32, knn_d2.f32, knn_found.i32, leaf_count.i32,
// box_count.i32, fcp_wide.i32, box_count_wide.i32, tri_ids.i32, tri_d2.f32,
// tri_P.f32, tri_ids_indexed.i32, tri_nodes.u8, tri_prims.u32}.  Exit
// status 0 on success.
#include "cuBQL/bvh.h"
#include "cuBQL/builder/cuda.h"
#include "cuBQL/queries/points/findClosest.h"
#include "cuBQL/queries/points/knn.h"
#include "cuBQL/queries/triangles/findClosest.h"
#include "cuBQL/traversal/fixedBoxQuery.h"

#include <cstdio>
#include <string>
#include <vector>

using namespace cuBQL;

constexpr int K = 8;

__global__ void computeBoxes(box3f *d_boxes, const vec3f *d_data, int numData) {
  int tid = threadIdx.x + blockIdx.x * blockDim.x;
  if (tid >= numData) return;

  d_boxes[tid] = box3f().including(d_data[tid]);
}

__global__ void runQueries(bvh3f bvh, const vec3f *d_data, const vec3f *d_queries, int numQueries,
                           int *d_closest) {
  int tid = threadIdx.x + blockIdx.x * blockDim.x;
  if (tid >= numQueries) return;

  vec3f queryPoint = d_queries[tid];
  int closestID = cuBQL::points::findClosest(/* the cubql bvh we've built */
                                             bvh,
                                             /* data that this bvh was built over*/
                                             d_data, queryPoint);
  d_closest[tid] = closestID;
}

__global__ void runKNN(bvh3f bvh, const vec3f *d_data, const vec3f *d_queries, int numQueries, int *ids,
                       float *d2, int *found) {
  int tid = threadIdx.x + blockIdx.x * blockDim.x;
  if (tid >= numQueries) return;
  knn::Candidate cand[K];
  knn::Result r = points::findKNN(cand, K, bvh, d_data, d_queries[tid]);
  for (int j = 0; j < K; j++) {
    ids[tid * K + j] = j < r.numFound ? cand[j].primID : -1;
    d2[tid * K + j] = j < r.numFound ? cand[j].sqrDist : INFINITY;
  }
  found[tid] = r.numFound;
}

// forEachLeaf with a fixed radius: number of points within it (the lambda
// never shrinks the radius)
__global__ void countInRadius(bvh3f bvh, const vec3f *d_data, const vec3f *d_queries, int numQueries,
                              float sqrRadius, int *count) {
  int tid = threadIdx.x + blockIdx.x * blockDim.x;
  if (tid >= numQueries) return;
  const vec3f q = d_queries[tid];
  int n = 0;
  auto leaf = [&](const uint32_t *ids, int cnt) -> float {
    for (int i = 0; i < cnt; i++)
      if (fSqrDistance_rd(d_data[ids[i]], q) <= sqrRadius) n++;
    return sqrRadius;
  };
  shrinkingRadiusQuery::forEachLeaf(leaf, bvh, q, sqrRadius);
  count[tid] = n;
}

// the wide form, as samples/s01_closestPoint_points_gpu/closestPoint_WideBVH.cu
enum { BVH_WIDTH = 8 };
typedef cuBQL::WideBVH<float, 3, BVH_WIDTH> wide_bvh3f;

__global__ void runQueriesWide(wide_bvh3f bvh, const vec3f *d_data, const vec3f *d_queries, int numQueries,
                               int *d_closest) {
  int tid = threadIdx.x + blockIdx.x * blockDim.x;
  if (tid >= numQueries) return;

  vec3f queryPoint = d_queries[tid];
  int closestID = cuBQL::points::findClosest(bvh, d_data, queryPoint);
  d_closest[tid] = closestID;
}

// fixedBoxQuery::forEachPrim: number of data points inside a box of half
// width h around each query (binary and wide trees give the same count)
template <typename BVH>
__global__ void countInBox(BVH bvh, const vec3f *d_data, const vec3f *d_queries, int numQueries, float h,
                           int *count) {
  int tid = threadIdx.x + blockIdx.x * blockDim.x;
  if (tid >= numQueries) return;
  const vec3f q = d_queries[tid];
  const box3f box(q - vec3f(h), q + vec3f(h));
  int n = 0;
  auto prim = [&](int primID) -> int {
    if (box.overlaps(box3f(d_data[primID]))) n++;
    return CUBQL_CONTINUE_TRAVERSAL;
  };
  fixedBoxQuery::forEachPrim(prim, bvh, box);
  count[tid] = n;
}

__global__ void computeTriBoxes(box3f *d_boxes, const Triangle *d_tris, int numTris) {
  int tid = threadIdx.x + blockIdx.x * blockDim.x;
  if (tid >= numTris) return;
  d_boxes[tid] = d_tris[tid].bounds();
}

__global__ void runTriQueries(bvh3f bvh, const Triangle *d_tris, const vec3f *d_queries, int numQueries,
                              int *ids, float *d2, vec3f *P) {
  int tid = threadIdx.x + blockIdx.x * blockDim.x;
  if (tid >= numQueries) return;
  triangles::FCPResult result;
  result.clear(INFINITY);
  triangles::findClosest(result, bvh, d_tris, d_queries[tid]);
  ids[tid] = result.primID;
  d2[tid] = result.sqrDist;
  P[tid] = result.P;
}

__global__ void runTriQueriesIndexed(bvh3f bvh, const vec3i *d_idx, const vec3f *d_verts,
                                     const vec3f *d_queries, int numQueries, int *ids) {
  int tid = threadIdx.x + blockIdx.x * blockDim.x;
  if (tid >= numQueries) return;
  triangles::FCPResult result;
  result.clear(INFINITY);
  triangles::fcp(result, d_queries[tid], bvh, d_idx, d_verts);
  ids[tid] = result.primID;
}

namespace {
template <typename T>
std::vector<T> load(const std::string &path) {
  std::vector<T> v;
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) return v;
  std::fseek(f, 0, SEEK_END);
  long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  v.resize(n / sizeof(T));
  if (std::fread(v.data(), sizeof(T), v.size(), f) != v.size()) v.clear();
  std::fclose(f);
  return v;
}
template <typename T>
bool save(const std::string &path, const T *d, size_t n) {
  std::vector<T> h(n);
  if (hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) return false;
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  bool ok = std::fwrite(h.data(), sizeof(T), n, f) == n;
  std::fclose(f);
  return ok;
}
template <typename T>
T *upload(const std::vector<T> &h) {
  T *d = nullptr;
  if (hipMalloc(&d, std::max<size_t>(1, h.size()) * sizeof(T)) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
  return d;
}
template <typename T>
T *alloc(size_t n) {
  T *d = nullptr;
  return hipMalloc(&d, std::max<size_t>(1, n) * sizeof(T)) == hipSuccess ? d : nullptr;
}
int fail(const char *what) {
  std::fprintf(stderr, "FAIL: %s\n", what);
  return 1;
}
}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) return fail("usage: device_api_sample <dir>");
  const std::string dir = argv[1];
  const auto data = load<vec3f>(dir + "/data.f32");
  const auto queries = load<vec3f>(dir + "/queries.f32");
  const auto verts = load<vec3f>(dir + "/verts.f32");
  const auto tris = load<vec3i>(dir + "/tris.i32");
  if (data.empty() || queries.empty() || verts.empty() || tris.empty()) return fail("missing input");
  const int numData = (int)data.size(), numQueries = (int)queries.size(), numTris = (int)tris.size();
  std::printf("data %d queries %d triangles %d\n", numData, numQueries, numTris);

  vec3f *d_data = upload(data), *d_queries = upload(queries), *d_verts = upload(verts);
  vec3i *d_idx = upload(tris);
  box3f *d_boxes = alloc<box3f>(numData);
  if (!d_data || !d_queries || !d_verts || !d_idx || !d_boxes) return fail("upload");
  computeBoxes<<<divRoundUp(numData, 128), 128>>>(d_boxes, d_data, numData);

  bvh3f bvh;
  cuBQL::gpuBuilder(bvh, d_boxes, numData, BuildConfig());
  int *d_fcp = alloc<int>(numQueries), *d_kid = alloc<int>(numQueries * K), *d_kf = alloc<int>(numQueries);
  int *d_cnt = alloc<int>(numQueries);
  float *d_kd = alloc<float>(numQueries * K);
  runQueries<<<divRoundUp(numQueries, 128), 128>>>(bvh, d_data, d_queries, numQueries, d_fcp);
  runKNN<<<divRoundUp(numQueries, 128), 128>>>(bvh, d_data, d_queries, numQueries, d_kid, d_kd, d_kf);
  countInRadius<<<divRoundUp(numQueries, 128), 128>>>(bvh, d_data, d_queries, numQueries, 0.01f, d_cnt);
  int *d_bc = alloc<int>(numQueries);
  countInBox<<<divRoundUp(numQueries, 128), 128>>>(bvh, d_data, d_queries, numQueries, 0.05f, d_bc);
  if (hipDeviceSynchronize() != hipSuccess) return fail("point kernels");
  if (!save(dir + "/box_count.i32", d_bc, numQueries)) return fail("save box counts");
  {
    wide_bvh3f wbvh;
    cuBQL::gpuBuilder(wbvh, d_boxes, numData, BuildConfig());
    int *d_wfcp = alloc<int>(numQueries), *d_wbc = alloc<int>(numQueries);
    runQueriesWide<<<divRoundUp(numQueries, 128), 128>>>(wbvh, d_data, d_queries, numQueries, d_wfcp);
    countInBox<<<divRoundUp(numQueries, 128), 128>>>(wbvh, d_data, d_queries, numQueries, 0.05f, d_wbc);
    if (hipDeviceSynchronize() != hipSuccess) return fail("wide kernels");
    if (!save(dir + "/fcp_wide.i32", d_wfcp, numQueries) || !save(dir + "/box_count_wide.i32", d_wbc, numQueries))
      return fail("save wide");
    cuda::free(wbvh);
  }
  if (!save(dir + "/fcp.i32", d_fcp, numQueries) || !save(dir + "/knn_ids.i32", d_kid, numQueries * K) ||
      !save(dir + "/knn_d2.f32", d_kd, numQueries * K) || !save(dir + "/knn_found.i32", d_kf, numQueries) ||
      !save(dir + "/leaf_count.i32", d_cnt, numQueries))
    return fail("save points");
  cuda::free(bvh);

  // triangles: the reference's non-indexed Triangle array and the indexed form
  std::vector<Triangle> tri(numTris);
  for (int i = 0; i < numTris; i++) tri[i] = Triangle{verts[tris[i].x], verts[tris[i].y], verts[tris[i].z]};
  Triangle *d_tri = upload(tri);
  box3f *d_tb = alloc<box3f>(numTris);
  if (!d_tri || !d_tb) return fail("upload triangles");
  computeTriBoxes<<<divRoundUp(numTris, 128), 128>>>(d_tb, d_tri, numTris);
  bvh3f tbvh;
  BuildConfig cfg;
  cfg.makeLeafThreshold = 4;
  cuBQL::gpuBuilder(tbvh, d_tb, numTris, cfg);
  int *d_tid = alloc<int>(numQueries), *d_tid2 = alloc<int>(numQueries);
  float *d_td = alloc<float>(numQueries);
  vec3f *d_P = alloc<vec3f>(numQueries);
  runTriQueries<<<divRoundUp(numQueries, 128), 128>>>(tbvh, d_tri, d_queries, numQueries, d_tid, d_td, d_P);
  runTriQueriesIndexed<<<divRoundUp(numQueries, 128), 128>>>(tbvh, d_idx, d_verts, d_queries, numQueries, d_tid2);
  if (hipDeviceSynchronize() != hipSuccess) return fail("triangle kernels");
  if (!save(dir + "/tri_ids.i32", d_tid, numQueries) || !save(dir + "/tri_d2.f32", d_td, numQueries) ||
      !save(dir + "/tri_P.f32", (const float *)d_P, 3 * (size_t)numQueries) ||
      !save(dir + "/tri_ids_indexed.i32", d_tid2, numQueries) ||
      !save(dir + "/tri_nodes.u8", (const uint8_t *)tbvh.nodes, 32 * (size_t)tbvh.numNodes) ||
      !save(dir + "/tri_prims.u32", tbvh.primIDs, tbvh.numPrims))
    return fail("save triangles");
  cuda::free(tbvh);
  std::printf("OK device api\n");
  return 0;
}
