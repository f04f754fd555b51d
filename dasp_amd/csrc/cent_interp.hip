// cent_interp.hip - interpolatePhase (cmod/src/centmodule.c:1628-1758) on the device.
//
// When the phase grid (m x m) is not nsubx*n wide the reference resamples, per sub-aperture,
// the points y + suby*n, x + subx*n of the (nsubx*n)^2 output grid that spans the same extent:
// scale = (m-1)/(nsubx*n-1), separable cubic convolution with the Catmull-Rom matrix
// {0,2,0,0; -1,0,1,0; 2,-5,4,-1; -1,3,-3,1}/2, rows/columns outside the grid replaced by the
// edge ones, and the last output row/column taken at t = 1 from the last three input samples.
// A sample depends only on its global output coordinate, so the whole output grid is resampled
// once per frame (one thread per output point, float arithmetic in the reference's order) and
// the image kernels then read it as an ordinary nsubx*n grid.
#include "cent_internal.h"

namespace dasp {

namespace {

__device__ __forceinline__ void interp_coord(int o, int tot, int m, float scale, int *i0, float *t) {
  float v = scale * (float)o;
  int i = (int)floorf(v);
  v -= (float)i;
  i--;
  if (o == tot - 1) {   // special case required
    v = 1.f;
    i = m - 3;
  }
  *i0 = i;
  *t = v;
}

__device__ __forceinline__ float fetch(const float *in, int m, int r, int c) {
  r = r < 0 ? 0 : (r >= m ? m - 1 : r);
  c = c < 0 ? 0 : (c >= m ? m - 1 : c);
  return in[(long)r * m + c];
}

// 0.5 * [1 t t^2 t^3] . M . f with M the Catmull-Rom matrix (times 2), as the reference's
// tmp/p products: tmp[k] = sum_j mx[4k+j] f[j], value = sum_k t[k] tmp[k].
__device__ __forceinline__ float cubic(const float (&t)[4], float f0, float f1, float f2, float f3) {
  const float m0 = 2.f * f1;
  const float m1 = -f0 + f2;
  const float m2 = 2.f * f0 - 5.f * f1 + 4.f * f2 - f3;
  const float m3 = -f0 + 3.f * f1 - 3.f * f2 + f3;
  return t[0] * m0 + t[1] * m1 + t[2] * m2 + t[3] * m3;
}

}  // namespace

__global__ __launch_bounds__(256) void interp_phase_kernel(InterpArgs a) {
  const long per = (long)a.tot * a.tot;
  const long total = per * a.nframes;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (long)gridDim.x * blockDim.x) {
    const int frame = (int)(k / per);
    const long q = k - frame * per;
    const int Y = (int)(q / a.tot), X = (int)(q - (long)Y * a.tot);
    const float *in = a.in + (long)frame * a.in_stride;
    int iny, inx;
    float ty, tx;
    interp_coord(Y, a.tot, a.m, a.scale, &iny, &ty);
    interp_coord(X, a.tot, a.m, a.scale, &inx, &tx);
    float t[4], tt[4];
    t[0] = 0.5f;
    t[1] = 0.5f * ty;
    t[2] = t[1] * ty;
    t[3] = t[2] * ty;
    tt[0] = 0.5f;
    tt[1] = 0.5f * tx;
    tt[2] = tt[1] * tx;
    tt[3] = tt[2] * tx;
    float p[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int c = inx + i;
      float f0, f1, f2, f3;
      if (iny < 0) {   // reuse the first row
        f1 = fetch(in, a.m, 0, c);
        f2 = fetch(in, a.m, 1, c);
        f3 = fetch(in, a.m, 2, c);
        f0 = f1;
      } else if (iny + 4 > a.m) {   // reuse the last row
        f0 = fetch(in, a.m, a.m - 3, c);
        f1 = fetch(in, a.m, a.m - 2, c);
        f2 = fetch(in, a.m, a.m - 1, c);
        f3 = f2;
      } else {
        f0 = fetch(in, a.m, iny, c);
        f1 = fetch(in, a.m, iny + 1, c);
        f2 = fetch(in, a.m, iny + 2, c);
        f3 = fetch(in, a.m, iny + 3, c);
      }
      p[i] = cubic(t, f0, f1, f2, f3);
    }
    a.out[(long)frame * per + q] = cubic(tt, p[0], p[1], p[2], p[3]);
  }
}

hipError_t launch_interp(const InterpArgs &a, hipStream_t st) {
  const long total = (long)a.tot * a.tot * a.nframes;
  if (total <= 0) return hipSuccess;
  long blocks = (total + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(interp_phase_kernel, dim3((unsigned)blocks), dim3(256), 0, st, a);
  return hipGetLastError();
}

}  // namespace dasp
