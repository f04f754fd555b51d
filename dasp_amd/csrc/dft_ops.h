// dft_ops.h - operand types and complex GEMM steps of the DFT-GEMM forms (cent_dft.hip,
// cent_dft32.hip): hi/lo splits of FP32 values into 16-bit matrix-core operands and the
// three-product complex products on v_mfma_f32_32x32x16_{f16,bf16}.
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

// Complex operand split into hi/lo parts of re and im (per work item).
template <typename T8>
struct COp {
  T8 rh, rl, ih, il;
  __device__ __forceinline__ void set(const float (&re)[8], const float (&im)[8]) {
    Split<T8>::run(re, rh, rl);
    Split<T8>::run(im, ih, il);
  }
};

// Constant DFT operand: also carries -im, so that the subtraction in re*re - im*im needs no
// per-item negated copy.
template <typename T8>
struct CConst {
  T8 rh, rl, ih, il, nih, nil;
  __device__ __forceinline__ void set(const float (&re)[8], const float (&im)[8]) {
    float nim[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) nim[t] = -im[t];
    Split<T8>::run(re, rh, rl);
    Split<T8>::run(im, ih, il);
    Split<T8>::run(nim, nih, nil);
  }
};

// A wave's 64 lanes of one constant operand in LDS, member by member (structure of arrays): a
// lane's 16-byte member sits 16 bytes after its neighbour's, so each ds_read_b128 is free of
// bank conflicts (an array of 96-byte CConst records puts every fourth lane on the same banks).
template <typename T8>
struct CConstLds {
  T8 rh[64], rl[64], ih[64], il[64], nih[64], nil[64];
  __device__ __forceinline__ void put(int l, const CConst<T8> &c) {
    rh[l] = c.rh;
    rl[l] = c.rl;
    ih[l] = c.ih;
    il[l] = c.il;
    nih[l] = c.nih;
    nil[l] = c.nil;
  }
  __device__ __forceinline__ CConst<T8> get(int l) const {
    CConst<T8> c;
    c.rh = rh[l];
    c.rl = rl[l];
    c.ih = ih[l];
    c.il = il[l];
    c.nih = nih[l];
    c.nil = nil[l];
    return c;
  }
};

// (A.re + i A.im)(B.re + i B.im) accumulated into (cr, ci), three products per real product.
// Stage 1: A varies per item, B = W^T is constant (negation on B).
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
