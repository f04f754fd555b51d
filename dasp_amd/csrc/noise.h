// noise.h - photon noise from a pre-drawn standard normal (device code of every image path).
//
// The reference draws Poisson variates inside the module (rk_poisson, cmod/src/centmodule.c:
// 760-782: multiplication method below lambda=10, PTRS rejection above) from per-thread PCG32
// streams, so its realisation depends on the thread count. Here every pixel consumes exactly one
// pre-drawn N(0,1) value z, which makes GPU and CPU results comparable value for value:
//   lambda <= 0      -> 0 (as rk_poisson)
//   0 < lambda < 10  -> exact inverse CDF of u = Phi(z) (pmf summed from k = 0, at most 40 terms)
//   lambda >= 10     -> third-order Cornish-Fisher quantile of the smoothed lattice variable
//                       Y = X + U(-1/2, 1/2) (cumulants lambda, lambda+1/12, lambda, lambda-1/120),
//                       k = max(ceil(lambda + s*w - 1/2), 0); total-variation distance to the
//                       Poisson pmf about 0.041/lambda.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace dasp {

// 1/(k+1) for the inverse-CDF recurrence, folded at compile time (correctly rounded)
static __constant__ const float kPoisInv[40] = {
    1.0f / 1,  1.0f / 2,  1.0f / 3,  1.0f / 4,  1.0f / 5,  1.0f / 6,  1.0f / 7,  1.0f / 8,
    1.0f / 9,  1.0f / 10, 1.0f / 11, 1.0f / 12, 1.0f / 13, 1.0f / 14, 1.0f / 15, 1.0f / 16,
    1.0f / 17, 1.0f / 18, 1.0f / 19, 1.0f / 20, 1.0f / 21, 1.0f / 22, 1.0f / 23, 1.0f / 24,
    1.0f / 25, 1.0f / 26, 1.0f / 27, 1.0f / 28, 1.0f / 29, 1.0f / 30, 1.0f / 31, 1.0f / 32,
    1.0f / 33, 1.0f / 34, 1.0f / 35, 1.0f / 36, 1.0f / 37, 1.0f / 38, 1.0f / 39, 1.0f / 40};

// Phi(z) = erfc(-z / sqrt 2) / 2 in a short branch-free form: with x = |z| / sqrt 2,
// t = 1 / (1 + x / 2) and s = 1 - t, erfc(x) = t exp(-x^2 + s Q(s)), Q a degree-7 polynomial
// fitted (least squares on x in [0, 6.5]) to log(erfcx(x) / t) / s, the factor s keeping the
// exact zero at x = 0 so that Horner's rule does not cancel. Absolute error of Phi below 1.4e-7
// for |z| <= 6 in FP32 (the correctly rounded value is within 4.3e-8); the pre-drawn normals
// reach |z| = 6.7 at most (Box-Muller from 32-bit uniforms), where the error is smaller still.
__device__ inline float normal_cdf(float z) {
  const float x = fabsf(z) * 0.70710678118654752f;
  const float t = __builtin_amdgcn_rcpf(fmaf(0.5f, x, 1.0f));
  const float s = 1.0f - t;
  float q = 1.424979736e-01f;
  q = fmaf(q, s, -2.659215828e-01f);
  q = fmaf(q, s, 2.444661891e-02f);
  q = fmaf(q, s, 3.886624346e-02f);
  q = fmaf(q, s, 1.938725000e-01f);
  q = fmaf(q, s, 1.616563901e-01f);
  q = fmaf(q, s, -3.032593895e-01f);
  q = fmaf(q, s, -1.256757958e+00f);
  const float h = 0.5f * t * __builtin_amdgcn_exp2f(fmaf(q, s, -x * x) * 1.4426950408889634f);
  return z < 0.f ? h : 1.0f - h;
}

// exp(-lambda) for 0 < lambda < 10 on the hardware exp2: the product lambda * log2(e) is carried
// with its rounding error (and the tail of log2(e)), which enters as exp2(a_lo) = 1 + a_lo ln 2.
__device__ inline float exp_neg(float lam) {
  const float L = 1.44269502162933349609375f, Llo = 1.925963e-08f;   // log2(e) = L + Llo
  const float ah = -lam * L;
  const float al = fmaf(-lam, Llo, fmaf(-lam, L, -ah));
  const float e = __builtin_amdgcn_exp2f(ah);
  return fmaf(e, al * 0.69314718055994531f, e);
}

// 0 < lambda < 10: the inverse CDF of u = Phi(z), searched from k = 0.
__device__ inline float poisson_small(float lam, float z) {
  const float u = normal_cdf(z);
  float p = exp_neg(lam);
  float F = p;
  // k = the number of leading steps i (from 1) at which u > F(i-1) and the sum still grows,
  // F(i) > F(i-1). Both conditions are monotone: once u <= F it stays so (F never decreases),
  // and the sum can stop growing only past the mode, where the terms keep shrinking. So k is
  // counted without branches: a lane's flag drops at its first failing step and freezes its
  // count, the step count is wave-uniform (1/(k+1) comes from the table through the scalar
  // cache) and the wave leaves the loop, checked every four steps, once no lane is searching.
  // The arithmetic of each step is that of the plain search: the same k, bit for bit.
  // The loop is unrolled whole: 1/(k+1) folds into the instructions as literal constants, and the
  // count is an integer (one add with the flag as carry).
  int k = 0;
  bool go = u > F;
#pragma unroll
  for (int i = 0; i < 40; i += 4) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const float pn = p * (lam * kPoisInv[i + t]);
      const float Fn = F + pn;
      go = go && (Fn > F);
      k += go ? 1 : 0;
      go = go && (u > Fn);
      p = pn;
      F = Fn;
    }
    if (!__any(go)) break;
  }
  return (float)k;
}

// lambda >= 10: the Cornish-Fisher quantile (hardware square root and reciprocal, 1 ulp each: a
// count can differ only where lambda + s w - 1/2 lies within a few ulp of an integer).
__device__ inline float poisson_large(float lam, float z) {
  const float s = __builtin_amdgcn_sqrtf(lam + 1.0f / 12.0f);
  const float s2 = s * s;
  const float r2 = __builtin_amdgcn_rcpf(s2);
  const float g1 = lam * (r2 * __builtin_amdgcn_rcpf(s));
  const float g2 = (lam - 1.0f / 120.0f) * (r2 * r2);
  const float z2 = z * z;
  const float w = z + g1 * (z2 - 1.0f) * (1.0f / 6.0f) + g2 * (z2 * z - 3.0f * z) * (1.0f / 24.0f) -
                  g1 * g1 * (2.0f * z2 * z - 5.0f * z) * (1.0f / 36.0f);
  const float k = ceilf(lam + s * w - 0.5f);
  return k > 0.f ? k : 0.f;
}

__device__ inline float poisson_from_normal(float lam, float z) {
  if (!(lam > 0.f)) return 0.f;
  if (lam < 10.f) return poisson_small(lam, z);
  return poisson_large(lam, z);
}

}  // namespace dasp
