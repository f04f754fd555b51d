// host_math.h - host-side numerics for the engine's set-up tables (double precision).
#pragma once
#include <complex>
#include <cstdint>
#include <vector>

namespace dasp {

using cd = std::complex<double>;

// In-place 1-D DFT of length n with stride; sign -1 forward, +1 inverse (unnormalised).
// Radix-2 for powers of two, direct summation otherwise.
void host_dft1(cd *x, int n, long stride, int sign);
// In-place 2-D DFT of an (m x n) row-major array.
void host_dft2(cd *x, int m, int n, int sign);

// flipArray (cmod/src/centmodule.c:275-349) on a square float array, mirror = 0 or 1.
void host_flip_array(float *a, int n, int mirror);

// Full complex spectrum of the half-complex (FFTW R2HC x R2HC) 2-D array a (m x n): the K with
// hcmul(R2HC(x), a) == R2HC-layout of DFT(x) * K for real x (util/correlation.py:97-137 for even
// sizes; odd sizes through the impulse response of the reference's product).
void host_hc_to_spectrum(const float *a, int m, int n, cd *K);

// First 5 rows of the inverse of the 6x6 normal matrix of [x^2, xy, y^2, x, y, 1] over a
// fitsize x fitsize grid (makeFitMx, cmod/src/centmodule.c:997-1057).
bool host_fit_matrix(int fitsize, float out[30]);

// PCG32 (O'Neill 2014, XSH-RR 64/32) and Box-Muller standard normals.
struct Pcg32 {
  uint64_t state = 0, inc = 1;
  void seed(uint64_t initstate, uint64_t initseq);
  uint32_t next();
  double uniform();   // in (0, 1)
};
void host_draw_normals(uint64_t seed, uint64_t stream, long n, float *out);

}  // namespace dasp
