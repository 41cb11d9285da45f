// io.hpp -- the reference's binary dump format for points and results.
//
// Counterpart of cuBQL::samples::loadBinary / saveBinary / convert
// (samples/common/IO.h:31-58, 60-66): a file holds a host size_t count N
// (8 bytes, little endian) followed by N raw items.  The reference's point
// files hold vec_t<double,D> items (testing/common/generatePoints.cpp:50-55)
// and its gold files float items (testing/points/closestPoint.h:150-160).
// Header-only host C++; no HIP types.
#pragma once
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace cuBQL {
namespace samples {

template <typename T>
std::vector<T> loadBinary(const std::string &fileName) {
  std::FILE *f = std::fopen(fileName.c_str(), "rb");
  if (!f) throw std::runtime_error("could not open '" + fileName + "'");
  uint64_t count = 0;
  if (std::fread(&count, sizeof(count), 1, f) != 1) {
    std::fclose(f);
    throw std::runtime_error("'" + fileName + "': no item count");
  }
  std::vector<T> data(count);
  const size_t got = count ? std::fread(data.data(), sizeof(T), count, f) : 0;
  std::fclose(f);
  if (got != count)
    throw std::runtime_error("'" + fileName + "': " + std::to_string(got) + " of " + std::to_string(count) +
                             " items present");
  return data;
}

template <typename T>
void saveBinary(const std::string &fileName, const std::vector<T> &data) {
  std::FILE *f = std::fopen(fileName.c_str(), "wb");
  if (!f) throw std::runtime_error("could not create '" + fileName + "'");
  const uint64_t count = data.size();
  bool ok = std::fwrite(&count, sizeof(count), 1, f) == 1;
  if (count) ok = ok && std::fwrite(data.data(), sizeof(T), count, f) == count;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) throw std::runtime_error("could not write '" + fileName + "'");
}

// Points as stored by the reference (three doubles each) converted to
// float32 per component with round-to-nearest, as convert<float,3> does.
struct vec3d_raw {
  double x, y, z;
};
static_assert(sizeof(vec3d_raw) == 24, "vec3d layout");

inline std::vector<float> loadPointsAsFloat3(const std::string &fileName) {
  const std::vector<vec3d_raw> in = loadBinary<vec3d_raw>(fileName);
  std::vector<float> out(3 * in.size());
  for (size_t i = 0; i < in.size(); i++) {
    out[3 * i + 0] = (float)in[i].x;
    out[3 * i + 1] = (float)in[i].y;
    out[3 * i + 2] = (float)in[i].z;
  }
  return out;
}

}  // namespace samples
}  // namespace cuBQL
