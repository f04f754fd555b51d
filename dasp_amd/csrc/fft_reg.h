// fft_reg.h - small complex FFTs held in one lane's registers, for CDNA (gfx950) kernels.
//
// Forward transform X[k] = sum_i x[i] exp(-2 pi i k i / N), unnormalised (the FFTW_FORWARD
// convention of the reference, cmod/src/centmodule.c:3105). Radix-2 decimation in time; every
// twiddle is a compile-time constant, the trivial ones (1, -1, -i, +i) are applied by renames and
// the others inside the butterfly as fused multiply-adds (butterfly below). The index
// permutations are register renames and cost nothing.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include <utility>

namespace dasp {

struct cf {
  float x, y;
};

__device__ __forceinline__ cf cadd(cf a, cf b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cf csub(cf a, cf b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cf cmul(cf a, cf b) {
  return {fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x)};
}
__device__ __forceinline__ float cabs2(cf a) { return fmaf(a.x, a.x, a.y * a.y); }

// ---- compile-time trigonometry (double Taylor series after reduction to [-pi, pi]) ----------
namespace ct {
constexpr double PI = 3.14159265358979323846264338327950288;
constexpr double reduce(double x) {
  while (x > PI) x -= 2 * PI;
  while (x < -PI) x += 2 * PI;
  return x;
}
constexpr double sin(double x) {
  x = reduce(x);
  double t = x, s = x;
  for (int k = 1; k < 40; ++k) {
    t *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
    s += t;
  }
  return s;
}
constexpr double cos(double x) {
  x = reduce(x);
  double t = 1.0, s = 1.0;
  for (int k = 1; k < 40; ++k) {
    t *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
    s += t;
  }
  return s;
}
}  // namespace ct

template <int B, int E, typename Fn>
__device__ __forceinline__ void static_for(Fn &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// Multiply a by exp(-2 pi i J / M) (J, M compile-time), exploiting the exact special values.
template <int J, int M>
__device__ __forceinline__ cf twiddle(cf a) {
  constexpr int j = ((J % M) + M) % M;
  if constexpr (j == 0) {
    return a;
  } else if constexpr (2 * j == M) {
    return {-a.x, -a.y};
  } else if constexpr (4 * j == M) {  // -i
    return {a.y, -a.x};
  } else if constexpr (4 * j == 3 * M) {  // +i
    return {-a.y, a.x};
  } else if constexpr (8 * j == M) {  // (1-i)/sqrt2
    constexpr float s = 0.70710678118654752440f;
    return {(a.x + a.y) * s, (a.y - a.x) * s};
  } else if constexpr (8 * j == 3 * M) {  // (-1-i)/sqrt2
    constexpr float s = 0.70710678118654752440f;
    return {(a.y - a.x) * s, -(a.x + a.y) * s};
  } else if constexpr (8 * j == 5 * M) {  // (-1+i)/sqrt2
    constexpr float s = 0.70710678118654752440f;
    return {-(a.x + a.y) * s, (a.x - a.y) * s};
  } else if constexpr (8 * j == 7 * M) {  // (1+i)/sqrt2
    constexpr float s = 0.70710678118654752440f;
    return {(a.x - a.y) * s, (a.x + a.y) * s};
  } else {
    constexpr float c = (float)ct::cos(-2.0 * ct::PI * j / M);
    constexpr float s = (float)ct::sin(-2.0 * ct::PI * j / M);
    return cmul(a, cf{c, s});
  }
}

constexpr double cabsd(double v) { return v < 0 ? -v : v; }

// Radix-2 butterfly u <- u + w t, t <- u - w t with w = exp(-2 pi i J / M). Trivial twiddles
// (1, -1, -i, +i) are exact renames; the others use the scaled rotation
//   w t = c ((t.x - g t.y) + i (t.y + g t.x)),  g = s / c     (|c| >= |s|)
//   w t = s ((h t.x - t.y) + i (h t.y + t.x)),  h = c / s     (|s| >  |c|)
// (w = c + i s), so that each output component is one fused multiply-add: six instructions per
// butterfly instead of eight (complex product, then add and subtract).
template <int J, int M>
__device__ __forceinline__ void butterfly(cf &u, cf &t) {
  constexpr int j = ((J % M) + M) % M;
  if constexpr (j == 0 || 2 * j == M || 4 * j == M || 4 * j == 3 * M) {
    const cf wt = twiddle<j, M>(t);
    const cf a = u;
    u = cadd(a, wt);
    t = csub(a, wt);
  } else {
    constexpr double th = -2.0 * ct::PI * j / M;
    constexpr double cd = ct::cos(th), sd = ct::sin(th);
    float vx, vy, m;
    if constexpr (cabsd(cd) >= cabsd(sd)) {
      constexpr float g = (float)(sd / cd);
      m = (float)cd;
      vx = fmaf(-g, t.y, t.x);
      vy = fmaf(g, t.x, t.y);
    } else {
      constexpr float h = (float)(cd / sd);
      m = (float)sd;
      vx = fmaf(h, t.x, -t.y);
      vy = fmaf(h, t.y, t.x);
    }
    const cf a = u;
    u = {fmaf(m, vx, a.x), fmaf(m, vy, a.y)};
    t = {fmaf(-m, vx, a.x), fmaf(-m, vy, a.y)};
  }
}

template <int N>
constexpr int ilog2() {
  int l = 0;
  while ((1 << l) < N) ++l;
  return l;
}
template <int BITS>
constexpr int bitrev(int v) {
  int r = 0;
  for (int b = 0; b < BITS; ++b) r |= ((v >> b) & 1) << (BITS - 1 - b);
  return r;
}

// Decimation-in-time stages S0..log2(N) on y (bit-reversed input order); stage s combines
// pairs of 2^(s-1)-point transforms into 2^s-point ones.
template <int N, int S0>
__device__ __forceinline__ void dit_stages(cf (&y)[N]) {
  constexpr int L = ilog2<N>();
  static_for<S0, L + 1>([&](auto s) {
    constexpr int m = 1 << decltype(s)::value;
    constexpr int h = m / 2;
    static_for<0, N / m>([&](auto g) {
      constexpr int k = decltype(g)::value * m;
      static_for<0, h>([&](auto jj) {
        constexpr int j = decltype(jj)::value;
        butterfly<j, m>(y[k + j], y[k + j + h]);
      });
    });
  });
}

// In-place forward FFT of length N (power of two, N <= 128) on x[0..N-1].
template <int N>
__device__ __forceinline__ void fft_fwd(cf (&x)[N]) {
  static_assert((N & (N - 1)) == 0 && N >= 1 && N <= 128, "power-of-two length");
  if constexpr (N == 1) {
    return;
  } else {
    constexpr int L = ilog2<N>();
    cf y[N];
    static_for<0, N>([&](auto i) { y[bitrev<L>(decltype(i)::value)] = x[decltype(i)::value]; });
    dit_stages<N, 1>(y);
    static_for<0, N>([&](auto i) { x[decltype(i)::value] = y[decltype(i)::value]; });
  }
}

// Zero-padded forward FFT by decimation in time: x holds the NP leading inputs of an F-point
// transform whose other inputs are zero (F = R*NP, R a power of two);
//   y[k] = sum_{i<NP} x[i] exp(-2 pi i k i / F),  k < F.
// In bit-reversed order the first log2(R) stages see one nonzero input per group of R, so after
// them each group holds that input R times and nothing is computed; the remaining stages run as
// in fft_fwd. Cheaper than R transforms of NP points on pre-twiddled inputs: the
// twiddles sit inside the scaled butterflies (324 instead of 352 instructions for 16 -> 32).
template <int NP, int F>
__device__ __forceinline__ void fft_fwd_zp(const cf (&x)[NP], cf (&y)[F]) {
  constexpr int R = F / NP, L = ilog2<F>(), LR = ilog2<R>();
  static_assert(R * NP == F && (R & (R - 1)) == 0 && (F & (F - 1)) == 0, "power-of-two lengths");
  static_for<0, NP>([&](auto gg) {
    constexpr int g = decltype(gg)::value;
    constexpr int i0 = bitrev<L>(g * R);   // < NP: the low log2(R) bits of g*R are zero
    static_for<0, R>([&](auto t) { y[g * R + decltype(t)::value] = x[i0]; });
  });
  dit_stages<F, LR + 1>(y);
}

// In-place unnormalised inverse FFT, X[k] = sum_i x[i] exp(+2 pi i k i / N) (the FFTW_BACKWARD
// convention): swapping real and imaginary parts turns the forward transform into the inverse,
// and the swaps are register renames.
template <int N>
__device__ __forceinline__ void fft_inv(cf (&x)[N]) {
  static_for<0, N>([&](auto i) { x[decltype(i)::value] = {x[decltype(i)::value].y, x[decltype(i)::value].x}; });
  fft_fwd<N>(x);
  static_for<0, N>([&](auto i) { x[decltype(i)::value] = {x[decltype(i)::value].y, x[decltype(i)::value].x}; });
}

}  // namespace dasp
