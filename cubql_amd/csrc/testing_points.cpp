// testing_points.cpp -- the reference's gold-file point tests on this
// library's GPU path.
//
// Same command lines, files and pass rule as the reference's
// testing/points/closestPoint.h:80-203 and testing/points/knn5.h:70-190:
//
//   cubql_testing_points closestPoint -d data.bin -q queries.bin -g gold.bin
//                        [-r|--rebuild-gold] [-p N] [--builder NAME]
//   cubql_testing_points knn5 ...
//
// data/queries: the reference's point dumps (size_t count + vec3d items,
// read with cuBQL::samples::loadBinary and converted to float); gold: one
// float per query, the distance to the closest point (closestPoint) or to
// the 5th nearest (knn5), INFINITY if none.  With --rebuild-gold the results
// are written to the gold file instead of compared.  Otherwise every result
// must be within 1e-3 of gold (closestPoint.h:48-53); the program prints the
// average result and the average and largest differences and exits 1 if any
// result is outside the threshold.  --builder picks the BVH builder: sm
// (gpuBuilder, the default), radix (63-bit Morton), radix30, rebin, sah.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "cubql_amd/cubql_capi.h"
#include "cubql_amd/io.hpp"

namespace {

void usage(const std::string &error) {
  if (!error.empty()) std::cout << "Error: " << error << "\n\n";
  std::cout << "Usage: ./cubql_testing_points closestPoint|knn5 -d dataPoints.bin -q queryPoints.bin "
               "-g goldResults.bin [--rebuild-gold] [-p numPrint] [--builder sm|radix|radix30|rebin|sah]"
            << std::endl;
  std::exit(error.empty() ? 0 : 1);
}

void check_hip(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    std::cerr << what << ": " << hipGetErrorString(e) << std::endl;
    std::exit(2);
  }
}

void check(int st, const char *what) {
  if (st != CUBQL_OK) {
    std::cerr << what << " failed (" << st << "): " << cubql_last_error_string() << std::endl;
    std::exit(2);
  }
}

template <typename T>
T *upload(const std::vector<T> &v) {
  T *d = nullptr;
  check_hip(hipMalloc((void **)&d, std::max<size_t>(1, v.size()) * sizeof(T)), "hipMalloc");
  if (!v.empty()) check_hip(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
  return d;
}

}  // namespace

int main(int ac, char **av) {
  if (ac < 2) usage("no test named (closestPoint or knn5)");
  const std::string test = av[1];
  if (test != "closestPoint" && test != "knn5") usage("unknown test '" + test + "'");
  std::string dataFile, queryFile, goldFile, builder = "sm";
  bool rebuildGold = false;
  int numPrint = 0;
  for (int i = 2; i < ac; i++) {
    const std::string a = av[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= ac) usage("missing value after '" + a + "'");
      return av[++i];
    };
    if (a == "-g")
      goldFile = next();
    else if (a == "-d")
      dataFile = next();
    else if (a == "-q")
      queryFile = next();
    else if (a == "-r" || a == "--rebuild-gold")
      rebuildGold = true;
    else if (a == "-p")
      numPrint = std::atoi(next().c_str());
    else if (a == "--builder")
      builder = next();
    else
      usage("un-recognized cmd-line argument '" + a + "'");
  }
  if (dataFile.empty()) usage("no data file name specified");
  if (queryFile.empty()) usage("no query file name specified");
  if (goldFile.empty()) usage("no gold file name specified");

  const std::vector<float> data = cuBQL::samples::loadPointsAsFloat3(dataFile);
  const std::vector<float> queries = cuBQL::samples::loadPointsAsFloat3(queryFile);
  const uint32_t numData = (uint32_t)(data.size() / 3), numQueries = (uint32_t)(queries.size() / 3);
  std::vector<cubql_box3f> boxes(numData);
  for (uint32_t i = 0; i < numData; i++)
    for (int d = 0; d < 3; d++) boxes[i].lower[d] = boxes[i].upper[d] = data[3 * i + d];
  float *dData = upload(data), *dQueries = upload(queries);
  cubql_box3f *dBoxes = upload(boxes);

  std::cout << "computing bvh (" << builder << ")" << std::endl;
  cubql_bvh3f bvh{nullptr, 0, nullptr, 0};
  cubql_build_config cfg{0, 8, CUBQL_SPATIAL_MEDIAN};
  if (builder == "sm")
    check(cubql_gpu_builder_f3(&bvh, dBoxes, numData, &cfg, nullptr, nullptr), "gpuBuilder");
  else if (builder == "radix")
    check(cubql_radix_builder_f3(&bvh, dBoxes, numData, &cfg, nullptr, nullptr), "radixBuilder");
  else if (builder == "radix30")
    check(cubql_radix30_builder_f3(&bvh, dBoxes, numData, &cfg, nullptr, nullptr), "radix30Builder");
  else if (builder == "rebin")
    check(cubql_rebin_radix_builder_f3(&bvh, dBoxes, numData, &cfg, nullptr, nullptr), "rebinRadixBuilder");
  else if (builder == "sah")
    check(cubql_sah_builder_f3(&bvh, dBoxes, numData, &cfg, 16, nullptr, nullptr), "sahBuilder");
  else
    usage("unknown builder '" + builder + "'");

  std::cout << "launching queries" << std::endl;
  const int k = test == "knn5" ? 5 : 1;
  int32_t *dIds = nullptr;
  float *dD2 = nullptr;
  check_hip(hipMalloc((void **)&dIds, (size_t)std::max(1u, numQueries) * k * sizeof(int32_t)), "hipMalloc");
  check_hip(hipMalloc((void **)&dD2, (size_t)std::max(1u, numQueries) * k * sizeof(float)), "hipMalloc");
  if (numQueries) {
    if (k == 1)
      check(cubql_fcp_points_f3(&bvh, dData, dQueries, numQueries, INFINITY, dIds, dD2, nullptr), "findClosest");
    else
      check(cubql_knn_points_f3(&bvh, dData, dQueries, numQueries, k, INFINITY, dIds, dD2, nullptr, nullptr),
            "findKNN");
  }
  std::vector<float> d2((size_t)numQueries * k);
  check_hip(hipDeviceSynchronize(), "queries");
  if (numQueries) check_hip(hipMemcpy(d2.data(), dD2, d2.size() * sizeof(float), hipMemcpyDeviceToHost), "download");
  // closestPoint: distance to the closest point (closestPoint.h:36-45);
  // knn5: sqrt of the 5th smallest squared distance (knn5.h:34-46)
  std::vector<float> results(numQueries);
  for (uint32_t i = 0; i < numQueries; i++) results[i] = std::sqrt(d2[(size_t)i * k + (k - 1)]);
  if (numPrint != 0)
    for (uint32_t i = 0; i < numQueries; i++)
      if (numPrint < 0 || (int)i < numPrint) std::cout << "[" << i << "] res = " << results[i] << std::endl;

  int rc = 0;
  if (rebuildGold) {
    std::cout << "asked to rebuild gold results: saving instead of comparing" << std::endl;
    cuBQL::samples::saveBinary(goldFile, results);
  } else {
    const std::vector<float> gold = cuBQL::samples::loadBinary<float>(goldFile);
    if (gold.size() != results.size()) {
      std::cout << "gold file has " << gold.size() << " results, expected " << results.size() << std::endl;
      rc = 1;
    } else {
      const double threshold = 1e-3;
      double avgGold = 0, avgDiff = 0, maxDiff = 0;
      int finiteGold = 0, finiteDiff = 0;
      for (float g : gold)
        if (std::isfinite(g)) {
          avgGold += g;
          finiteGold++;
        }
      for (size_t i = 0; i < results.size(); i++) {
        const double diff = std::fabs((double)results[i] - (double)gold[i]);
        if (std::isfinite(diff)) {
          maxDiff = std::max(maxDiff, diff);
          avgDiff += diff;
          finiteDiff++;
        }
        // a NaN difference (inf - inf) passes, as in the reference
        if (diff >= threshold) {
          std::cout << "!! different result for index " << i << ": ours says " << results[i] << ", gold says "
                    << gold[i] << ", a difference of " << diff << std::endl;
          rc = 1;
        }
      }
      std::cout << "avg result                   " << avgGold / std::max(1, finiteGold) << std::endl;
      std::cout << "avg difference (rel to gold) " << avgDiff / std::max(1, finiteDiff) << std::endl;
      std::cout << "max difference (rel to gold) " << maxDiff << std::endl;
      std::cout << "(tolerance threshold " << threshold << ")" << std::endl;
      std::cout << (rc ? "FAILED" : "PASSED") << std::endl;
    }
  }
  check_hip(hipFree(dIds), "hipFree");
  check_hip(hipFree(dD2), "hipFree");
  check_hip(hipFree(dData), "hipFree");
  check_hip(hipFree(dQueries), "hipFree");
  check_hip(hipFree(dBoxes), "hipFree");
  check(cubql_free_f3(&bvh, nullptr, nullptr), "free");
  return rc;
}
