// host_math.cpp - host-side numerics for the engine's set-up tables.
#include "host_math.h"

#include <cmath>
#include <cstring>

namespace dasp {

static const double kPi = 3.14159265358979323846264338327950288;

void host_dft1(cd *x, int n, long stride, int sign) {
  if (n <= 1) return;
  if ((n & (n - 1)) == 0) {
    for (int i = 1, j = 0; i < n; ++i) {
      int bit = n >> 1;
      for (; j & bit; bit >>= 1) j ^= bit;
      j ^= bit;
      if (i < j) std::swap(x[i * stride], x[j * stride]);
    }
    for (int len = 2; len <= n; len <<= 1) {
      const double ang = sign * 2 * kPi / len;
      for (int i = 0; i < n; i += len)
        for (int k = 0; k < len / 2; ++k) {
          const cd w(std::cos(ang * k), std::sin(ang * k));
          cd u = x[(i + k) * stride], v = x[(i + k + len / 2) * stride] * w;
          x[(i + k) * stride] = u + v;
          x[(i + k + len / 2) * stride] = u - v;
        }
    }
    return;
  }
  std::vector<cd> tmp(n);
  for (int k = 0; k < n; ++k) {
    cd acc = 0;
    for (int j = 0; j < n; ++j) {
      const double ang = sign * 2 * kPi * (double)((long)j * k % n) / n;
      acc += x[j * stride] * cd(std::cos(ang), std::sin(ang));
    }
    tmp[k] = acc;
  }
  for (int k = 0; k < n; ++k) x[k * stride] = tmp[k];
}

void host_dft2(cd *x, int m, int n, int sign) {
  for (int r = 0; r < m; ++r) host_dft1(x + (long)r * n, n, 1, sign);
  for (int c = 0; c < n; ++c) host_dft1(x + c, m, n, sign);
}

static void flip_quad(float *a, int n, int r0, int c0, int h, int w) {
  for (int i = 0; i < h / 2; ++i)
    for (int j = 0; j < w; ++j)
      std::swap(a[(r0 + i) * n + c0 + j], a[(r0 + h - 1 - i) * n + c0 + w - 1 - j]);
  if (h % 2) {
    const int i = h / 2;
    for (int j = 0; j < w / 2; ++j)
      std::swap(a[(r0 + i) * n + c0 + j], a[(r0 + h - 1 - i) * n + c0 + w - 1 - j]);
  }
}

void host_flip_array(float *a, int n, int mirror) {
  const int w1 = n / 2, w3 = (n + 1) / 2;
  if (mirror == 0 || n % 2) {
    flip_quad(a, n, 0, 0, w1, w1);
    flip_quad(a, n, w1, 0, w3, w1);
    flip_quad(a, n, 0, w1, w1, w3);
    flip_quad(a, n, w1, w1, w3, w3);
    if (mirror) flip_quad(a, n, 0, 0, n, n);
    return;
  }
  std::vector<float> t(w1);
  for (int i = 0; i < w1; ++i) {
    std::memcpy(t.data(), &a[i * n], sizeof(float) * w1);
    std::memcpy(&a[i * n], &a[(i + w1) * n + w1], sizeof(float) * w1);
    std::memcpy(&a[(i + w1) * n + w1], t.data(), sizeof(float) * w1);
    std::memcpy(t.data(), &a[i * n + w1], sizeof(float) * w1);
    std::memcpy(&a[i * n + w1], &a[(i + w1) * n], sizeof(float) * w1);
    std::memcpy(&a[(i + w1) * n], t.data(), sizeof(float) * w1);
  }
}

// The half-complex product of calcCorrelation (cmod/src/centmodule.c:1436-1482), in double.
static void hc_multiply(double *a, const double *B, int m, int n) {
  auto A = [&](int i, int j) -> double & { return a[i * n + j]; };
  auto b = [&](int i, int j) { return B[i * n + j]; };
  const bool neven = n % 2 == 0, meven = m % 2 == 0;
  A(0, 0) *= b(0, 0);
  if (neven) A(0, n / 2) *= b(0, n / 2);
  if (meven) {
    A(m / 2, 0) *= b(m / 2, 0);
    if (neven) A(m / 2, n / 2) *= b(m / 2, n / 2);
  }
  for (int i = 1; i < (n + 1) / 2; ++i) {
    double r1 = A(0, i) * b(0, i) - A(0, n - i) * b(0, n - i);
    double r2 = A(0, i) * b(0, n - i) + A(0, n - i) * b(0, i);
    A(0, i) = r1;
    A(0, n - i) = r2;
    if (meven) {
      double r3 = A(m / 2, i) * b(m / 2, i) - A(m / 2, n - i) * b(m / 2, n - i);
      double r4 = A(m / 2, i) * b(m / 2, n - i) + A(m / 2, n - i) * b(m / 2, i);
      A(m / 2, i) = r3;
      A(m / 2, n - i) = r4;
    }
  }
  for (int i = 1; i < (m + 1) / 2; ++i) {
    double r5 = A(i, 0) * b(i, 0) - A(m - i, 0) * b(m - i, 0);
    double r6 = A(i, 0) * b(m - i, 0) + A(m - i, 0) * b(i, 0);
    A(i, 0) = r5;
    A(m - i, 0) = r6;
    if (neven) {
      double r7 = A(i, n / 2) * b(i, n / 2) - A(m - i, n / 2) * b(m - i, n / 2);
      double r8 = A(i, n / 2) * b(m - i, n / 2) + A(m - i, n / 2) * b(i, n / 2);
      A(i, n / 2) = r7;
      A(m - i, n / 2) = r8;
    }
    for (int j = 1; j < (n + 1) / 2; ++j) {
      double aij = A(i, j), amn = A(m - i, n - j), ain = A(i, n - j), amj = A(m - i, j);
      double r1 = aij * b(i, j) + amn * b(m - i, n - j) - ain * b(i, n - j) - amj * b(m - i, j);
      double r2 = aij * b(m - i, n - j) + amn * b(i, j) + amj * b(i, n - j) + ain * b(m - i, j);
      double r3 = aij * b(i, n - j) - amn * b(m - i, j) + ain * b(i, j) - amj * b(m - i, n - j);
      double r4 = aij * b(m - i, j) - amn * b(i, n - j) + amj * b(i, j) - ain * b(m - i, n - j);
      A(i, j) = r1;
      A(m - i, n - j) = r2;
      A(i, n - j) = r3;
      A(m - i, j) = r4;
    }
  }
}

// R2HC along one axis of a real array (FFTW half-complex layout), double.
static void r2hc_axis(double *a, int m, int n, int axis) {
  const int len = axis == 1 ? n : m, cnt = axis == 1 ? m : n;
  std::vector<cd> t(len);
  for (int c = 0; c < cnt; ++c) {
    for (int k = 0; k < len; ++k) t[k] = axis == 1 ? a[c * n + k] : a[k * n + c];
    host_dft1(t.data(), len, 1, -1);
    for (int k = 0; k < len; ++k) {
      double v = k <= len / 2 ? t[k].real() : t[len - k].imag();
      if (axis == 1) a[c * n + k] = v;
      else a[k * n + c] = v;
    }
  }
}

// HC2R (unnormalised) along one axis.
static void hc2r_axis(double *a, int m, int n, int axis) {
  const int len = axis == 1 ? n : m, cnt = axis == 1 ? m : n;
  std::vector<cd> t(len);
  for (int c = 0; c < cnt; ++c) {
    auto at = [&](int k) { return axis == 1 ? a[c * n + k] : a[k * n + c]; };
    for (int k = 0; k < len; ++k) {
      if (k == 0 || (len % 2 == 0 && k == len / 2)) t[k] = cd(at(k), 0);
      else if (k < (len + 1) / 2) t[k] = cd(at(k), at(len - k));
      else t[k] = std::conj(cd(at(len - k), at(k)));
    }
    host_dft1(t.data(), len, 1, +1);
    for (int k = 0; k < len; ++k) {
      if (axis == 1) a[c * n + k] = t[k].real();
      else a[k * n + c] = t[k].real();
    }
  }
}

void host_hc_to_spectrum(const float *pat, int m, int n, cd *K) {
  std::vector<double> B(pat, pat + (long)m * n);
  if (m % 2 == 0 && n % 2 == 0) {
    // unpack (util/correlation.py:97-137) for even sizes
    auto a = [&](int i, int j) { return B[i * n + j]; };
    for (long q = 0; q < (long)m * n; ++q) K[q] = 0;
    auto k = [&](int i, int j) -> cd & { return K[i * n + j]; };
    for (int i = 1; i < (m + 1) / 2; ++i) {
      k(i, 0) = cd(a(i, 0), a(m - i, 0));
      k(m - i, 0) = cd(a(i, 0), -a(m - i, 0));
      k(i, n / 2) = cd(a(i, n / 2), a(m - i, n / 2));
      k(m - i, n / 2) = cd(a(i, n / 2), -a(m - i, n / 2));
      for (int j = 1; j < (n + 1) / 2; ++j) {
        if (i == 1) {
          k(0, j) = cd(a(0, j), a(0, n - j));
          k(0, n - j) = cd(a(0, j), -a(0, n - j));
          k(m / 2, j) = cd(a(m / 2, j), a(m / 2, n - j));
          k(m / 2, n - j) = cd(a(m / 2, j), -a(m / 2, n - j));
        }
        k(i, j) = cd(a(i, j) - a(m - i, n - j), a(i, n - j) + a(m - i, j));
        k(m - i, n - j) = cd(a(i, j) - a(m - i, n - j), -(a(i, n - j) + a(m - i, j)));
        k(i, n - j) = cd(a(i, j) + a(m - i, n - j), a(m - i, j) - a(i, n - j));
        k(m - i, j) = cd(a(i, j) + a(m - i, n - j), -(a(m - i, j) - a(i, n - j)));
      }
    }
    k(0, 0) = cd(a(0, 0), 0);
    k(0, n / 2) = cd(a(0, n / 2), 0);
    k(m / 2, 0) = cd(a(m / 2, 0), 0);
    k(m / 2, n / 2) = cd(a(m / 2, n / 2), 0);
    return;
  }
  // general sizes: impulse response of R2HC -> hc_multiply -> HC2R, then its DFT / (m n)
  std::vector<double> d((long)m * n, 0.0);
  d[0] = 1.0;
  r2hc_axis(d.data(), m, n, 1);
  r2hc_axis(d.data(), m, n, 0);
  hc_multiply(d.data(), B.data(), m, n);
  hc2r_axis(d.data(), m, n, 1);
  hc2r_axis(d.data(), m, n, 0);
  for (long q = 0; q < (long)m * n; ++q) K[q] = cd(d[q] / ((double)m * n), 0);
  host_dft2(K, m, n, -1);
}

bool host_fit_matrix(int fitsize, float out[30]) {
  double A[6][12];
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 12; ++c) A[r][c] = (c - 6 == r) ? 1.0 : 0.0;
  for (int y = 0; y < fitsize; ++y)
    for (int x = 0; x < fitsize; ++x) {
      const double v[6] = {(double)x * x, (double)x * y, (double)y * y, (double)x, (double)y, 1.0};
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) A[r][c] += v[r] * v[c];
    }
  for (int col = 0; col < 6; ++col) {   // Gauss-Jordan with partial pivoting
    int piv = col;
    for (int r = col + 1; r < 6; ++r)
      if (std::fabs(A[r][col]) > std::fabs(A[piv][col])) piv = r;
    if (std::fabs(A[piv][col]) < 1e-300) return false;
    if (piv != col)
      for (int c = 0; c < 12; ++c) std::swap(A[piv][c], A[col][c]);
    const double inv = 1.0 / A[col][col];
    for (int c = 0; c < 12; ++c) A[col][c] *= inv;
    for (int r = 0; r < 6; ++r)
      if (r != col) {
        const double f = A[r][col];
        if (f != 0)
          for (int c = 0; c < 12; ++c) A[r][c] -= f * A[col][c];
      }
  }
  for (int r = 0; r < 5; ++r)
    for (int c = 0; c < 6; ++c) out[r * 6 + c] = (float)A[r][6 + c];
  return true;
}

void Pcg32::seed(uint64_t initstate, uint64_t initseq) {
  state = 0u;
  inc = (initseq << 1u) | 1u;
  next();
  state += initstate;
  next();
}

uint32_t Pcg32::next() {
  const uint64_t old = state;
  state = old * 6364136223846793005ULL + inc;
  const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
  const uint32_t rot = (uint32_t)(old >> 59u);
  return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

double Pcg32::uniform() { return ((double)next() + 0.5) * (1.0 / 4294967296.0); }

void host_draw_normals(uint64_t seed, uint64_t stream, long n, float *out) {
  Pcg32 g;
  g.seed(seed, stream);
  long i = 0;
  while (i < n) {
    const double u1 = g.uniform(), u2 = g.uniform();
    const double r = std::sqrt(-2.0 * std::log(u1));
    const double th = 2 * kPi * u2;
    out[i++] = (float)(r * std::cos(th));
    if (i < n) out[i++] = (float)(r * std::sin(th));
  }
}

}  // namespace dasp
