// dft_ops.h - operand types and complex GEMM steps of the DFT-GEMM forms (cent_dft.hip,
// cent_dft32.hip): hi/lo splits of FP32 values into 16-bit matrix-core operands and the
// three-product complex products on v_mfma_f32_32x32x16_{f16,bf16} (COMP = true, the compensated
// forms), or the single-product ones on the hi parts alone (COMP = false: one rounding of each
// operand to 16 bits, a quarter of the matrix-core work and no lo parts to form).
//
// MFMA 32x32x16 lane layout (checked on the device): A[m][kk] sits in lane m + 32*(kk/8), slot
// kk%8; B[kk][nn] in lane nn + 32*(kk/8), slot kk%8; D[m][nn] in lane nn + 32*((m/4)%2),
// register 4*(m/8) + m%4.
#pragma once
#include "cent_internal.h"

namespace dasp {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

// hi/lo split of 8 floats into two 16-bit vectors (the residual x - hi is exact in FP32).
template <typename T8>
struct Split;

template <>
struct Split<h8> {
  // hi: round-toward-zero pair conversion (one v_cvt_pkrtz per two values) stored straight into
  // the operand's dwords; lo = x - hi is exact in FP32 and is rounded to FP16 by a mixed-precision
  // FMA, fma(hi, -1, x) (v_fma_mix, one instruction per value)
  static __device__ __forceinline__ void run(const float (&x)[8], h8 &hi, h8 &lo) {
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    u4 hw, lw;
    // an opaque 1.0 (one SGPR): fma(x, 1, -hi) would otherwise be folded into x - hi, which is
    // then emitted as two conversions, a subtraction and a repack instead of one v_fma_mix
    float one = 1.0f;
    asm("" : "+s"(one));
#pragma unroll
    for (int t = 0; t < 8; t += 2) {
      const h2 hp = __builtin_bit_cast(h2, __builtin_amdgcn_cvt_pkrtz(x[t], x[t + 1]));
      h2 lp;
      lp[0] = (_Float16)__builtin_fmaf(x[t], one, -(float)hp[0]);
      lp[1] = (_Float16)__builtin_fmaf(x[t + 1], one, -(float)hp[1]);
      hw[t / 2] = __builtin_bit_cast(unsigned, hp);
      lw[t / 2] = __builtin_bit_cast(unsigned, lp);
    }
    hi = __builtin_bit_cast(h8, hw);
    lo = __builtin_bit_cast(h8, lw);
  }
};

template <>
struct Split<b8> {
  static __device__ __forceinline__ void run(const float (&x)[8], b8 &hi, b8 &lo) {
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const __bf16 hv = (__bf16)x[t];
      hi[t] = hv;
      lo[t] = (__bf16)(x[t] - (float)hv);
    }
  }
};

__device__ __forceinline__ f16v mfma(const h8 &a, const h8 &b, const f16v &c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f16v mfma(const b8 &a, const b8 &b, const f16v &c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// The 16-bit value nearest below in magnitude (FP16: one round-toward-zero pair conversion per
// two values) or nearest (BF16) of 8 floats: the single-product forms' operands.
template <typename T8>
struct Round;

template <>
struct Round<h8> {
  static __device__ __forceinline__ void run(const float (&x)[8], h8 &hi) {
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    u4 hw;
#pragma unroll
    for (int t = 0; t < 8; t += 2) hw[t / 2] = __builtin_bit_cast(unsigned, __builtin_amdgcn_cvt_pkrtz(x[t], x[t + 1]));
    hi = __builtin_bit_cast(h8, hw);
  }
};

template <>
struct Round<b8> {
  static __device__ __forceinline__ void run(const float (&x)[8], b8 &hi) {
#pragma unroll
    for (int t = 0; t < 8; ++t) hi[t] = (__bf16)x[t];
  }
};

template <typename T8>
struct Elem;
template <>
struct Elem<h8> {
  using type = _Float16;
};
template <>
struct Elem<b8> {
  using type = __bf16;
};

// Complex operand split into hi/lo parts of re and im (per work item); without COMP only the hi
// parts are formed (rl and il stay unused).
template <typename T8, bool COMP = true>
struct COp {
  T8 rh, rl, ih, il;
  __device__ __forceinline__ void set(const float (&re)[8], const float (&im)[8]) {
    if constexpr (COMP) {
      Split<T8>::run(re, rh, rl);
      Split<T8>::run(im, ih, il);
    } else {
      Round<T8>::run(re, rh);
      Round<T8>::run(im, ih);
    }
  }
};

// Constant DFT operand: also carries -im, so that the subtraction in re*re - im*im needs no
// per-item negated copy. The constants are built once per wave; without COMP only the hi parts.
template <typename T8, bool COMP = true>
struct CConst {
  T8 rh, rl, ih, il, nih, nil;
  __device__ __forceinline__ void set(const float (&re)[8], const float (&im)[8]) {
    float nim[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) nim[t] = -im[t];
    if constexpr (COMP) {
      Split<T8>::run(re, rh, rl);
      Split<T8>::run(im, ih, il);
      Split<T8>::run(nim, nih, nil);
    } else {
      // nearest 16-bit values of the constants (built once, so the exact rounding costs nothing)
      using E = typename Elem<T8>::type;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        rh[t] = (E)re[t];
        ih[t] = (E)im[t];
        nih[t] = (E)nim[t];
      }
    }
  }
};

// A wave's 64 lanes of one constant operand in LDS, member by member (structure of arrays): a
// lane's 16-byte member sits 16 bytes after its neighbour's, so each ds_read_b128 is free of
// bank conflicts (an array of 96-byte CConst records puts every fourth lane on the same banks).
// Without COMP the lo members hold one unused entry.
template <typename T8, bool COMP = true>
struct CConstLds {
  static constexpr int L = COMP ? 64 : 1;
  T8 rh[64], ih[64], nih[64], rl[L], il[L], nil[L];
  __device__ __forceinline__ void put(int l, const CConst<T8, COMP> &c) {
    rh[l] = c.rh;
    ih[l] = c.ih;
    nih[l] = c.nih;
    if constexpr (COMP) {
      rl[l] = c.rl;
      il[l] = c.il;
      nil[l] = c.nil;
    }
  }
  __device__ __forceinline__ CConst<T8, COMP> get(int l) const {
    CConst<T8, COMP> c;
    c.rh = rh[l];
    c.ih = ih[l];
    c.nih = nih[l];
    if constexpr (COMP) {
      c.rl = rl[l];
      c.il = il[l];
      c.nil = nil[l];
    }
    return c;
  }
};

// (A.re + i A.im)(B.re + i B.im) accumulated into (cr, ci), three products per real product
// (COMP) or one.
// Stage 1: A varies per item, B = W^T is constant (negation on B).
template <typename T8>
__device__ __forceinline__ void cgemm_vc(const COp<T8, false> &A, const CConst<T8, false> &B, f16v &cr, f16v &ci) {
  cr = mfma(A.rh, B.rh, cr);
  ci = mfma(A.rh, B.ih, ci);
  cr = mfma(A.ih, B.nih, cr);
  ci = mfma(A.ih, B.rh, ci);
}
template <typename T8>
__device__ __forceinline__ void cgemm_cv(const CConst<T8, false> &A, const COp<T8, false> &B, f16v &cr, f16v &ci) {
  cr = mfma(A.rh, B.rh, cr);
  ci = mfma(A.rh, B.ih, ci);
  cr = mfma(A.nih, B.ih, cr);
  ci = mfma(A.ih, B.rh, ci);
}
template <typename T8>
__device__ __forceinline__ void cgemm_vc(const COp<T8> &A, const CConst<T8> &B, f16v &cr, f16v &ci) {
  cr = mfma(A.rh, B.rh, cr);
  ci = mfma(A.rh, B.ih, ci);
  cr = mfma(A.rh, B.rl, cr);
  ci = mfma(A.rh, B.il, ci);
  cr = mfma(A.rl, B.rh, cr);
  ci = mfma(A.rl, B.ih, ci);
  cr = mfma(A.ih, B.nih, cr);
  ci = mfma(A.ih, B.rh, ci);
  cr = mfma(A.ih, B.nil, cr);
  ci = mfma(A.ih, B.rl, ci);
  cr = mfma(A.il, B.nih, cr);
  ci = mfma(A.il, B.rh, ci);
}

// Stage 2: A = W is constant (negation on A), B varies per item.
template <typename T8>
__device__ __forceinline__ void cgemm_cv(const CConst<T8> &A, const COp<T8> &B, f16v &cr, f16v &ci) {
  cr = mfma(A.rh, B.rh, cr);
  ci = mfma(A.rh, B.ih, ci);
  cr = mfma(A.rh, B.rl, cr);
  ci = mfma(A.rh, B.il, ci);
  cr = mfma(A.rl, B.rh, cr);
  ci = mfma(A.rl, B.ih, ci);
  cr = mfma(A.nih, B.ih, cr);
  ci = mfma(A.ih, B.rh, ci);
  cr = mfma(A.nih, B.il, cr);
  ci = mfma(A.ih, B.rl, ci);
  cr = mfma(A.nil, B.ih, cr);
  ci = mfma(A.il, B.rh, ci);
}

template <bool FAST>
__device__ __forceinline__ void dft_sincos(float th, float *sn, float *cs) {
  phase_sincos<FAST>(th, sn, cs);
}

}  // namespace dasp
