// build_common.h -- device helpers shared by the builders.
#pragma once
#include "common.h"

namespace cubql {

__device__ inline bool box_is_valid(const cubql_box3f &b) { return b.lower[0] <= b.upper[0]; }

__device__ inline cubql_box3f load_box(const cubql_box3f *p, uint32_t i) {
  const float2 *f = reinterpret_cast<const float2 *>(p + i);
  float2 a = f[0], b = f[1], c = f[2];
  cubql_box3f r;
  r.lower[0] = a.x; r.lower[1] = a.y; r.lower[2] = b.x;
  r.upper[0] = b.y; r.upper[1] = c.x; r.upper[2] = c.y;
  return r;
}

__device__ inline float coord(const float4 &v, int d) {
  return d == 0 ? v.x : (d == 1 ? v.y : v.z);
}

// Split choice (sm_builder.h:194-220 / spatialMedian.h:90-91): widest axis
// by strict '>' (ties to the lower axis), pos = midpoint.
__device__ inline void choose_split(const float lo[3], const float hi[3], int &dim, float &pos) {
  float s0 = hi[0] - lo[0], s1 = hi[1] - lo[1], s2 = hi[2] - lo[2];
  dim = s2 > fmaxf(s0, s1) ? 2 : (s1 > s0 ? 1 : 0);
  float l = dim == 0 ? lo[0] : (dim == 1 ? lo[1] : lo[2]);
  float h = dim == 0 ? hi[0] : (dim == 1 ? hi[1] : hi[2]);
  pos = (l + h) / 2.f;
}

// block-wide min/max reduction of 12 floats (left lo/hi, right lo/hi)
template <int BLOCK>
__device__ inline void block_reduce_bounds(float (&v)[12], float *smem /*12*BLOCK/64*/) {
  constexpr int NW = BLOCK / kWave;
  const uint32_t lane = lane_id(), wid = threadIdx.x / kWave;
#pragma unroll
  for (int j = 0; j < 12; j++) {
    bool isMin = (j % 6) < 3;
    v[j] = isMin ? wave_min(v[j]) : wave_max(v[j]);
  }
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < 12; j++) smem[j * NW + wid] = v[j];
  }
  __syncthreads();
  if (threadIdx.x < 12) {
    int j = threadIdx.x;
    bool isMin = (j % 6) < 3;
    float r = smem[j * NW];
    for (int w = 1; w < NW; w++) r = isMin ? fminf(r, smem[j * NW + w]) : fmaxf(r, smem[j * NW + w]);
    smem[j * NW] = r;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 12; j++) v[j] = smem[j * NW];
  __syncthreads();
}

}  // namespace cubql
