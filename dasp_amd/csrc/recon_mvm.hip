// recon_mvm.hip - the reconstructor's matrix-vector step on the device, fed by the engine's slopes.
//
// science/tomoRecon.py (reconType spmxSVD / spmxGI / svd / pinv / reg with a dense reconmx) gathers
// the slope vector data of length ncents (the x slopes of every active sub-aperture of every WFS,
// then the y slopes: centIndex, :131-150; getInput :682-700), then
//   tmp = -data . reconmx            (:1140-1142; reconmx (ncents, nmodes), or its transpose :1146)
//   outputData = decay * outputData + gains * tmp[:nacts]      (decay :1090-1107, :1174).
// Here the gather, the product and the integrator run on the device for nframes independent slope
// vectors; the matrix is read once per launch of up to kMaxF frames.
//
// Layout 0 (reconmx rows = slopes, row stride ld >= nacts): a split-K product over 256-column x
// kReconRows-row tiles (partial sums per tile to a scratch); a second kernel adds the tiles of a
// column in a fixed order (the result is reproducible run to run) and applies the integrator.
// Roofline: HBM, 4 * ncents * nacts bytes.
// Layout 1 (reconmx rows = actuators, row stride ld >= ncents): one wave per actuator row, lanes
// stride the row (coalesced, float4), a wave sum per frame, and the integrator in the same kernel.
#include <hip/hip_runtime.h>

#include "recon_internal.h"

namespace dasp {

constexpr int kCols = 256;                    // columns per block: 64 lanes x 4
constexpr int kWaveRows = kReconRows / 4;     // rows per wave of a block's tile

// Block (cb, rc): a 256-column x kReconRows-row tile, wave w taking rows [w*kWaveRows, (w+1)*
// kWaveRows) of it, each lane four columns (float4 loads when rows are 16-byte aligned), eight rows
// of loads in flight. The slopes of a row come from the gathered copy at a block-uniform address
// (scalar loads). The four waves' sums meet in LDS in a fixed order; wave 0 writes the tile's
// partial sums (chunk, frame, column).
template <int F, bool VEC>
__global__ __launch_bounds__(256) void recon_rows_kernel(ReconArgs a) {
  __shared__ float red[3][F][4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r0 = blockIdx.y * kReconRows + w * kWaveRows;
  const int nr = max(0, min(kWaveRows, a.ncents - r0));
  const int c0 = blockIdx.x * kCols + 4 * lane;
  float acc[F][4];
#pragma unroll
  for (int f = 0; f < F; ++f)
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[f][k] = 0.f;
  if (c0 < a.nacts) {
    const float *dat = a.gathered + r0;
    const float *row = a.mx + (long)r0 * a.ld + c0;
    if (VEC && c0 + 3 < a.nacts) {
      int r = 0;
      for (; r + 8 <= nr; r += 8) {
        float4 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = *(const float4 *)(row + (long)(r + u) * a.ld);
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
          for (int f = 0; f < F; ++f) {
            const float s = dat[(long)f * a.gstride + r + u];
            acc[f][0] = fmaf(s, v[u].x, acc[f][0]);
            acc[f][1] = fmaf(s, v[u].y, acc[f][1]);
            acc[f][2] = fmaf(s, v[u].z, acc[f][2]);
            acc[f][3] = fmaf(s, v[u].w, acc[f][3]);
          }
      }
      for (; r < nr; ++r) {
        const float4 v = *(const float4 *)(row + (long)r * a.ld);
#pragma unroll
        for (int f = 0; f < F; ++f) {
          const float s = dat[(long)f * a.gstride + r];
          acc[f][0] = fmaf(s, v.x, acc[f][0]);
          acc[f][1] = fmaf(s, v.y, acc[f][1]);
          acc[f][2] = fmaf(s, v.z, acc[f][2]);
          acc[f][3] = fmaf(s, v.w, acc[f][3]);
        }
      }
    } else {
      const int nc = min(4, a.nacts - c0);
      for (int r = 0; r < nr; ++r) {
        const float *p = row + (long)r * a.ld;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const float v = k < nc ? p[k] : 0.f;
#pragma unroll
          for (int f = 0; f < F; ++f) acc[f][k] = fmaf(dat[(long)f * a.gstride + r], v, acc[f][k]);
        }
      }
    }
  }
  if (w > 0) {
#pragma unroll
    for (int f = 0; f < F; ++f)
#pragma unroll
      for (int k = 0; k < 4; ++k) red[w - 1][f][k][lane] = acc[f][k];
  }
  __syncthreads();
  if (w == 0 && c0 < a.nacts) {
#pragma unroll
    for (int f = 0; f < F; ++f) {
      float *o = a.part + ((long)blockIdx.y * F + f) * a.nacts + c0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float s = ((acc[f][k] + red[0][f][k][lane]) + red[1][f][k][lane]) + red[2][f][k][lane];
        if (c0 + k < a.nacts) o[k] = s;
      }
    }
  }
}

// Chunks summed (wave w takes every fourth chunk from w, in order; the four sums are added in a
// fixed order), then acts = decay * acts - gain * sum. Block: 64 columns of one frame.
__global__ __launch_bounds__(256) void recon_reduce_kernel(ReconArgs a, int F, int nchunks) {
  __shared__ float red[3][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + lane, f = blockIdx.y;
  float s = 0.f;
  if (c < a.nacts)
    for (int k = w; k < nchunks; k += 4) s += a.part[((long)k * F + f) * a.nacts + c];
  if (w > 0) red[w - 1][lane] = s;
  __syncthreads();
  if (w == 0 && c < a.nacts) {
    s = ((s + red[0][lane]) + red[1][lane]) + red[2][lane];
    float *o = a.acts + (long)f * a.acts_stride + c;
    const float dec = a.decay ? a.decay[c] : a.decay_s;
    const float g = a.gains ? a.gains[c] : a.gain_s;
    *o = dec * *o - g * s;
  }
}

// The slope vectors in the reconstructor's order (frames x gstride, gstride = ncents rounded up to
// a multiple of 4 so that every frame starts 16-byte aligned), read by both layouts.
__global__ void recon_gather_kernel(ReconArgs a, int F) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)F * a.ncents) return;
  const int f = (int)(t / a.ncents), c = (int)(t - (long)f * a.ncents);
  a.gathered[(long)f * a.gstride + c] = a.slopes[(long)f * a.slope_stride + a.index[c]];
}

// One wave per actuator row; float4 loads of the row and of the gathered slopes when rows are
// 16-byte aligned, scalar loads for the tail.
template <int F, bool VEC>
__global__ __launch_bounds__(256) void recon_cols_kernel(ReconArgs a) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (row >= a.nacts) return;
  const float *m = a.mx + (long)row * a.ld;
  float acc[F];
#pragma unroll
  for (int f = 0; f < F; ++f) acc[f] = 0.f;
  int c = lane;
  if (VEC) {
    const int nv = a.ncents & ~3;
    for (int c4 = 4 * lane; c4 < nv; c4 += 256) {
      const float4 v = *(const float4 *)(m + c4);
#pragma unroll
      for (int f = 0; f < F; ++f) {
        const float4 g = *(const float4 *)(a.gathered + (long)f * a.gstride + c4);
        acc[f] = fmaf(v.x, g.x, fmaf(v.y, g.y, fmaf(v.z, g.z, fmaf(v.w, g.w, acc[f]))));
      }
    }
    c = nv + lane;
  }
  for (; c < a.ncents; c += 64) {
    const float v = m[c];
#pragma unroll
    for (int f = 0; f < F; ++f) acc[f] = fmaf(v, a.gathered[(long)f * a.gstride + c], acc[f]);
  }
#pragma unroll
  for (int f = 0; f < F; ++f) {
    float s = acc[f];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) {
      float *o2 = a.acts + (long)f * a.acts_stride + row;
      const float dec = a.decay ? a.decay[row] : a.decay_s;
      const float g = a.gains ? a.gains[row] : a.gain_s;
      *o2 = dec * *o2 - g * s;
    }
  }
}

template <int F>
static hipError_t launch_f(const ReconArgs &a, int layout, hipStream_t st) {
  const long ng = (long)F * a.ncents;
  hipLaunchKernelGGL(recon_gather_kernel, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, st, a, F);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (layout == 0) {
    const int nchunks = (a.ncents + kReconRows - 1) / kReconRows;
    const dim3 grid((a.nacts + kCols - 1) / kCols, nchunks);
      const bool vec = a.ld % 4 == 0 && ((unsigned long long)a.mx & 15) == 0;
    if (vec) hipLaunchKernelGGL((recon_rows_kernel<F, true>), grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL((recon_rows_kernel<F, false>), grid, dim3(256), 0, st, a);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(recon_reduce_kernel, dim3((unsigned)((a.nacts + 63) / 64), F), dim3(256), 0, st, a, F, nchunks);
    return hipGetLastError();
  }
  const bool vec = a.ld % 4 == 0 && ((unsigned long long)a.mx & 15) == 0;
  if (vec) hipLaunchKernelGGL((recon_cols_kernel<F, true>), dim3((unsigned)((a.nacts + 3) / 4)), dim3(256), 0, st, a);
  else hipLaunchKernelGGL((recon_cols_kernel<F, false>), dim3((unsigned)((a.nacts + 3) / 4)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_recon(const ReconArgs &a, int layout, int F, hipStream_t st) {
  switch (F) {
    case 1: return launch_f<1>(a, layout, st);
    case 2: return launch_f<2>(a, layout, st);
    case 3: return launch_f<3>(a, layout, st);
    case 4: return launch_f<4>(a, layout, st);
    case 5: return launch_f<5>(a, layout, st);
    case 6: return launch_f<6>(a, layout, st);
    case 7: return launch_f<7>(a, layout, st);
    case 8: return launch_f<8>(a, layout, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace dasp
