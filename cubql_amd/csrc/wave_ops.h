// wave_ops.h -- wave64 reductions with DPP and in-register lane permutes.
//
// GFX9/CDNA data-parallel primitives: row_shr 1,2,4,8 build per-row (16
// lanes) prefixes, row_bcast:15 / row_bcast:31 combine the four rows, the
// result is read from lane 63.  All 64 lanes must be active; callers give
// idle lanes the identity value.
#pragma once
#include "common.h"

namespace cubql {

template <int CTRL, int ROWMASK = 0xF, int BANKMASK = 0xF>
__device__ __forceinline__ float dpp_f(float old, float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), CTRL,
                                                    ROWMASK, BANKMASK, false));
}

__device__ __forceinline__ float wave_min_dpp(float v) {
  const float I = INFINITY;
  v = fminf(v, dpp_f<0x111>(I, v));        // row_shr:1
  v = fminf(v, dpp_f<0x112>(I, v));        // row_shr:2
  v = fminf(v, dpp_f<0x114>(I, v));        // row_shr:4
  v = fminf(v, dpp_f<0x118>(I, v));        // row_shr:8
  v = fminf(v, dpp_f<0x142, 0xA>(I, v));   // row_bcast:15 into rows 1,3
  v = fminf(v, dpp_f<0x143, 0xC>(I, v));   // row_bcast:31 into rows 2,3
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ __forceinline__ float wave_max_dpp(float v) {
  const float I = -INFINITY;
  v = fmaxf(v, dpp_f<0x111>(I, v));
  v = fmaxf(v, dpp_f<0x112>(I, v));
  v = fmaxf(v, dpp_f<0x114>(I, v));
  v = fmaxf(v, dpp_f<0x118>(I, v));
  v = fmaxf(v, dpp_f<0x142, 0xA>(I, v));
  v = fmaxf(v, dpp_f<0x143, 0xC>(I, v));
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Inclusive prefix scans with DPP (row_shr 1,2,4,8 inside rows, then
// row_bcast 15, and 31 unless the two 32-lane halves scan independently),
// on encoded bounds (enc_f keeps float order, so integer min/max equal the
// float ones) and counts: with an identity as the old value each step is
// one fused v_min/v_max/v_add_*_dpp.
template <int OP, bool CROSS32>  // OP: 0 min, 1 max, 2 add
__device__ __forceinline__ int32_t dpp_scan_i(int32_t v) {
  const int32_t I = OP == 0 ? INT32_MAX : (OP == 1 ? INT32_MIN : 0);
#define CUBQL_DPP_STEP(CTRL, ROWMASK)                                                        \
  {                                                                                          \
    const int32_t t = __builtin_amdgcn_update_dpp(I, v, CTRL, ROWMASK, 0xF, false);          \
    v = OP == 0 ? min(v, t) : (OP == 1 ? max(v, t) : v + t);                                 \
  }
  CUBQL_DPP_STEP(0x111, 0xF)
  CUBQL_DPP_STEP(0x112, 0xF)
  CUBQL_DPP_STEP(0x114, 0xF)
  CUBQL_DPP_STEP(0x118, 0xF)
  CUBQL_DPP_STEP(0x142, 0xA)
  if (CROSS32) CUBQL_DPP_STEP(0x143, 0xC)
#undef CUBQL_DPP_STEP
  return v;
}

// Wave-wide min/max of encoded bounds (enc_f), result in every lane.
__device__ __forceinline__ int32_t wave_min_i32(int32_t v) {
  return __builtin_amdgcn_readlane(dpp_scan_i<0, true>(v), 63);
}
__device__ __forceinline__ int32_t wave_max_i32(int32_t v) {
  return __builtin_amdgcn_readlane(dpp_scan_i<1, true>(v), 63);
}

// enc_f of x with -0 read as +0: integer order then equals float order,
// equality included, for every non-NaN x
__device__ __forceinline__ int32_t enc_cmp(float x) { return enc_f(x + 0.f); }

// push `v` to lane `dst` (a permutation of the 64 lanes)
__device__ __forceinline__ float lane_push_f(uint32_t dst, float v) {
  return __int_as_float(__builtin_amdgcn_ds_permute((int)(dst << 2), __float_as_int(v)));
}
__device__ __forceinline__ uint32_t lane_push_u(uint32_t dst, uint32_t v) {
  return (uint32_t)__builtin_amdgcn_ds_permute((int)(dst << 2), (int)v);
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const uint32_t l = __lane_id();
  return l ? (~0ull >> (64 - l)) : 0ull;
}

}  // namespace cubql
