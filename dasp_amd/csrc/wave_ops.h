// wave_ops.h - cross-lane sums and maxima on DPP (data-parallel primitives) for wave64.
//
// A DPP source operand costs nothing beyond the VALU instruction it feeds, where __shfl_xor goes
// through ds_bpermute (an LDS round trip plus index arithmetic). Group reductions over W
// consecutive lanes, result in every lane of the group:
//   W = 2  quad_perm [1,0,3,2]      W = 4  quad_perm [2,3,0,1]
//   W = 8  row_half_mirror (lane i <-> 7-i in each half row, after the quad step)
//   W = 16 row_mirror      (lane i <-> 15-i in each row, after the half-row step)
//   W = 32, 64: one ds_bpermute exchange per doubling (the rows are combined across)
// wave_total() reduces all 64 lanes with row_bcast15/row_bcast31 and returns the total from
// lane 63 as a wave-uniform value.
#pragma once
#include <hip/hip_runtime.h>

namespace dasp {

namespace dpp {
constexpr int kQuadXor1 = 0xB1;       // quad_perm [1,0,3,2]
constexpr int kQuadXor2 = 0x4E;       // quad_perm [2,3,0,1]
constexpr int kRowMirror = 0x140;
constexpr int kRowHalfMirror = 0x141;
constexpr int kRowBcast15 = 0x142;
constexpr int kRowBcast31 = 0x143;
}  // namespace dpp

template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}

// src moved by CTRL in the enabled rows, `old` in the others
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_update(float old, float src) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                               __builtin_bit_cast(int, src), CTRL, ROWS, 0xF,
                                                               false));
}

template <int W>
__device__ __forceinline__ float group_sum(float v) {
  static_assert(W >= 1 && W <= 64 && (W & (W - 1)) == 0, "power-of-two group");
  if constexpr (W >= 2) v += dpp_mov<dpp::kQuadXor1>(v);
  if constexpr (W >= 4) v += dpp_mov<dpp::kQuadXor2>(v);
  if constexpr (W >= 8) v += dpp_mov<dpp::kRowHalfMirror>(v);
  if constexpr (W >= 16) v += dpp_mov<dpp::kRowMirror>(v);
  if constexpr (W >= 32) v += __shfl_xor(v, 16);
  if constexpr (W >= 64) v += __shfl_xor(v, 32);
  return v;
}

template <int W>
__device__ __forceinline__ float group_max(float v) {
  static_assert(W >= 1 && W <= 64 && (W & (W - 1)) == 0, "power-of-two group");
  if constexpr (W >= 2) v = fmaxf(v, dpp_mov<dpp::kQuadXor1>(v));
  if constexpr (W >= 4) v = fmaxf(v, dpp_mov<dpp::kQuadXor2>(v));
  if constexpr (W >= 8) v = fmaxf(v, dpp_mov<dpp::kRowHalfMirror>(v));
  if constexpr (W >= 16) v = fmaxf(v, dpp_mov<dpp::kRowMirror>(v));
  if constexpr (W >= 32) v = fmaxf(v, __shfl_xor(v, 16));
  if constexpr (W >= 64) v = fmaxf(v, __shfl_xor(v, 32));
  return v;
}

// Sum over all 64 lanes, returned wave-uniform (called from whole-wave code).
__device__ __forceinline__ float wave_total(float v) {
  v = group_sum<16>(v);                                      // row totals in every lane
  v += dpp_update<dpp::kRowBcast15, 0xA>(0.f, v);            // rows 1, 3: + row 0, row 2
  v += dpp_update<dpp::kRowBcast31, 0xC>(0.f, v);            // rows 2, 3: + lane 31 (rows 0 + 1)
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}

// Maximum over all 64 lanes, wave-uniform.
__device__ __forceinline__ float wave_maximum(float v) {
  v = group_max<16>(v);
  const int b0 = __builtin_bit_cast(int, v);
  const float r0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(b0, 0));
  const float r1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(b0, 16));
  const float r2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(b0, 32));
  const float r3 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(b0, 48));
  return fmaxf(fmaxf(r0, r1), fmaxf(r2, r3));
}

}  // namespace dasp
