// cuBQL/math/vec.h -- vec3f / vec3i / box3f in the reference's spelling
// (math/vec.h, math/box.h) for the float, 3-D instantiation this library
// builds: the same member names (x, y, z; lower, upper), the same byte
// layout (12-byte vectors, 24-byte boxes) and the operations the query
// headers and typical user kernels use.  Host and device code.
#pragma once
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include "cubql_amd/cubql_capi.h"

#define CUBQL_BOTH __host__ __device__ inline
// no FMA contraction inside a function body: float results round as written
#define CUBQL_NOFMA _Pragma("clang fp contract(off)")

namespace cuBQL {

struct vec3f {
  float x, y, z;
  vec3f() = default;
  CUBQL_BOTH explicit vec3f(float v) : x(v), y(v), z(v) {}
  CUBQL_BOTH vec3f(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
  CUBQL_BOTH float operator[](int i) const { return i == 2 ? z : (i ? y : x); }
  CUBQL_BOTH float &operator[](int i) { return i == 2 ? z : (i ? y : x); }
};
struct vec3i {
  int x, y, z;
  vec3i() = default;
  CUBQL_BOTH vec3i(int x_, int y_, int z_) : x(x_), y(y_), z(z_) {}
  CUBQL_BOTH int operator[](int i) const { return i == 2 ? z : (i ? y : x); }
  CUBQL_BOTH int &operator[](int i) { return i == 2 ? z : (i ? y : x); }
};
static_assert(sizeof(vec3f) == 12 && sizeof(vec3i) == 12, "vec3 layout");

// Element-wise arithmetic without FMA contraction (vec.h:406-493).
CUBQL_BOTH vec3f operator+(vec3f a, vec3f b) { CUBQL_NOFMA return {a.x + b.x, a.y + b.y, a.z + b.z}; }
CUBQL_BOTH vec3f operator-(vec3f a, vec3f b) { CUBQL_NOFMA return {a.x - b.x, a.y - b.y, a.z - b.z}; }
CUBQL_BOTH vec3f operator*(vec3f a, vec3f b) { CUBQL_NOFMA return {a.x * b.x, a.y * b.y, a.z * b.z}; }
CUBQL_BOTH vec3f operator*(float s, vec3f a) { CUBQL_NOFMA return {s * a.x, s * a.y, s * a.z}; }
CUBQL_BOTH vec3f operator*(vec3f a, float s) { CUBQL_NOFMA return {a.x * s, a.y * s, a.z * s}; }
CUBQL_BOTH vec3f operator/(vec3f a, float s) { CUBQL_NOFMA return {a.x / s, a.y / s, a.z / s}; }
CUBQL_BOTH vec3f operator-(vec3f a) { return {-a.x, -a.y, -a.z}; }
CUBQL_BOTH bool operator==(vec3f a, vec3f b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
CUBQL_BOTH bool operator!=(vec3f a, vec3f b) { return !(a == b); }
CUBQL_BOTH vec3f min(vec3f a, vec3f b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
CUBQL_BOTH vec3f max(vec3f a, vec3f b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
CUBQL_BOTH float dot(vec3f a, vec3f b) { CUBQL_NOFMA return a.x * b.x + a.y * b.y + a.z * b.z; }
CUBQL_BOTH vec3f cross(vec3f a, vec3f b) { CUBQL_NOFMA
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
CUBQL_BOTH float sqrLength(vec3f v) { return dot(v, v); }
CUBQL_BOTH float length(vec3f v) { return sqrtf(dot(v, v)); }
CUBQL_BOTH vec3f normalize(vec3f v) { return v * (1.f / sqrtf(dot(v, v))); }
CUBQL_BOTH float clamp(float v, float lo = 0.f, float hi = 1.f) { return fminf(fmaxf(v, lo), hi); }
CUBQL_BOTH float reduce_max(vec3f v) { return fmaxf(fmaxf(v.x, v.y), v.z); }
CUBQL_BOTH float reduce_min(vec3f v) { return fminf(fminf(v.x, v.y), v.z); }
// squared distance (x*x + y*y) + z*z, as the traversal computes it
// (conservativeDistances.h:89-90, 99-100)
CUBQL_BOTH float fSqrDistance_rd(vec3f a, vec3f b) { return sqrLength(a - b); }

/*! box_t<float,3> (box.h:60-219): lower then upper; the default box is
    empty (lower = +inf, upper = -inf); lower.x > upper.x marks an invalid
    (ignored) primitive for the builders. */
struct box3f {
  vec3f lower, upper;
  CUBQL_BOTH box3f() : lower(INFINITY), upper(-INFINITY) {}
  CUBQL_BOTH explicit box3f(vec3f p) : lower(p), upper(p) {}
  CUBQL_BOTH box3f(vec3f lo, vec3f hi) : lower(lo), upper(hi) {}
  CUBQL_BOTH box3f including(vec3f p) const { return box3f(cuBQL::min(lower, p), cuBQL::max(upper, p)); }
  CUBQL_BOTH box3f including(const box3f &b) const {
    return box3f(cuBQL::min(lower, b.lower), cuBQL::max(upper, b.upper));
  }
  CUBQL_BOTH box3f &grow(vec3f p) { return *this = including(p); }
  CUBQL_BOTH box3f &grow(const box3f &b) { return *this = including(b); }
  CUBQL_BOTH box3f &extend(vec3f p) { return grow(p); }
  CUBQL_BOTH box3f &extend(const box3f &b) { return grow(b); }
  CUBQL_BOTH box3f &set_empty() { return *this = box3f(); }
  CUBQL_BOTH void clear() { set_empty(); }
  CUBQL_BOTH bool empty() const { return lower.x > upper.x; }
  CUBQL_BOTH vec3f center() const { return (lower + upper) / 2.f; }
  CUBQL_BOTH vec3f size() const { return upper - lower; }
  CUBQL_BOTH float get_lower(int d) const { return lower[d]; }
  CUBQL_BOTH float get_upper(int d) const { return upper[d]; }
  CUBQL_BOTH bool overlaps(const box3f &b) const {
    return lower.x <= b.upper.x && b.lower.x <= upper.x && lower.y <= b.upper.y &&
           b.lower.y <= upper.y && lower.z <= b.upper.z && b.lower.z <= upper.z;
  }
};
static_assert(sizeof(box3f) == sizeof(cubql_box3f), "box3f layout");

/*! squared distance from a point to a box (0 inside), clamp-then-dot as
    in shrinkingRadiusQuery.h:253-259 */
CUBQL_BOTH float fSqrDistance_rd(vec3f p, const box3f &b) {
  return fSqrDistance_rd(cuBQL::min(cuBQL::max(p, b.lower), b.upper), p);
}

CUBQL_BOTH int32_t divRoundUp(int32_t a, int32_t b) { return (a + b - 1) / b; }
CUBQL_BOTH uint32_t divRoundUp(uint32_t a, uint32_t b) { return (a + b - 1) / b; }
CUBQL_BOTH int64_t divRoundUp(int64_t a, int64_t b) { return (a + b - 1) / b; }
CUBQL_BOTH uint64_t divRoundUp(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

}  // namespace cuBQL
