// cuBQL/queries/triangles/Triangle.h -- Triangle (queries/triangles/
// Triangle.h:26-48): three vertices, bounds(), sample(u, v), area().
#pragma once
#include "cuBQL/bvh.h"

namespace cuBQL {

struct Triangle {
  CUBQL_BOTH box3f bounds() const { return box3f().including(a).including(b).including(c); }
  CUBQL_BOTH vec3f sample(float u, float v) const {
    if (u + v >= 1.f) { u = 1.f - u; v = 1.f - v; }
    return (1.f - u - v) * a + u * b + v * c;
  }
  vec3f a, b, c;
};

CUBQL_BOTH float area(Triangle tri) { return length(cross(tri.b - tri.a, tri.c - tri.a)); }

}  // namespace cuBQL
