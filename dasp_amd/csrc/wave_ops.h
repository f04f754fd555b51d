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

// Lane-half and row exchanges of gfx950 (v_permlane32_swap_b32: lanes 32-63 of the first
// register trade places with lanes 0-31 of the second; v_permlane16_swap_b32: rows 1 and 3 of the
// first with rows 0 and 2 of the second). Written as inline asm: the compiler's builtins for them
// return the first register for both results in this toolchain. The s_nop gives the swap the
// wait states it needs after the VALU instructions that wrote its operands.
__device__ __forceinline__ float halves_sum(float a, float b) {
  // lanes 0-31: a[l] + a[l + 32]; lanes 32-63: b[l - 32] + b[l]
  asm("s_nop 1\n\tv_permlane32_swap_b32 %0, %1" : "+v"(a), "+v"(b));
  return a + b;
}
__device__ __forceinline__ float rows_sum(float a, float b) {
  // row 0: a.r0 + a.r1, row 1: b.r0 + b.r1, row 2: a.r2 + a.r3, row 3: b.r2 + b.r3
  asm("s_nop 1\n\tv_permlane16_swap_b32 %0, %1" : "+v"(a), "+v"(b));
  return a + b;
}

// Totals over all 64 lanes of six values (a0..a2, b0..b2), wave-uniform, by a transposing
// reduction: two lane-half and row exchange levels pack the partial sums of several values into
// one register (one row per value) before the in-row DPP sums, 24 vector instructions against
// 54 for six wave_total calls.
__device__ __forceinline__ void wave_totals6(float (&a)[3], float (&b)[3]) {
  const float p0 = halves_sum(a[0], b[0]);   // lanes 0-31: a0, 32-63: b0 (32-lane partials)
  const float p1 = halves_sum(a[1], b[1]);
  const float p2 = halves_sum(a[2], b[2]);
  float q = rows_sum(p0, p1);                // rows: a0, a1, b0, b1 (16-lane partials)
  float r = rows_sum(p2, p2);                // rows: a2, a2, b2, b2
  q = group_sum<16>(q);
  r = group_sum<16>(r);
  const int qi = __builtin_bit_cast(int, q), ri = __builtin_bit_cast(int, r);
  a[0] = __builtin_bit_cast(float, __builtin_amdgcn_readlane(qi, 0));
  a[1] = __builtin_bit_cast(float, __builtin_amdgcn_readlane(qi, 16));
  b[0] = __builtin_bit_cast(float, __builtin_amdgcn_readlane(qi, 32));
  b[1] = __builtin_bit_cast(float, __builtin_amdgcn_readlane(qi, 48));
  a[2] = __builtin_bit_cast(float, __builtin_amdgcn_readlane(ri, 0));
  b[2] = __builtin_bit_cast(float, __builtin_amdgcn_readlane(ri, 32));
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
