// cent_noise.hip - the readout noise streams drawn on the device.
//
// The same standard normals as the host generator (host_math.cpp host_draw_normals): PCG32
// (O'Neill 2014, XSH-RR 64/32) seeded with (seed, stream), uniforms u = (x + 0.5) / 2^32 in
// double, Box-Muller pairs r cos(2 pi u2), r sin(2 pi u2) with r = sqrt(-2 ln u1), rounded to
// float. Pair p uses outputs 2p and 2p+1 of the stream, so a thread that starts at pair p jumps
// the generator ahead by 2p steps (LCG jump-ahead by squaring, O(log p)) and the result does
// not depend on how the pairs are split over threads. The integer sequence is identical to the
// host's; the double-precision log/sqrt/cos/sin are the device's, so a value can differ from the
// host's in its last float bit where the two libraries' doubles straddle a rounding boundary.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cent_internal.h"

namespace dasp {

namespace {

constexpr uint64_t kPcgMult = 6364136223846793005ULL;
constexpr int kPairsPerThread = 16;

struct DevPcg32 {
  uint64_t state, inc;
  __device__ void seed(uint64_t initstate, uint64_t initseq) {
    inc = (initseq << 1u) | 1u;
    state = inc;               // state 0 stepped once
    state += initstate;
    state = state * kPcgMult + inc;
  }
  // state after `delta` further steps: s -> A s + C with (A, C) composed by squaring
  __device__ void advance(uint64_t delta) {
    uint64_t acc_mult = 1u, acc_plus = 0u, cur_mult = kPcgMult, cur_plus = inc;
    while (delta > 0) {
      if (delta & 1u) {
        acc_mult *= cur_mult;
        acc_plus = acc_plus * cur_mult + cur_plus;
      }
      cur_plus = (cur_mult + 1u) * cur_plus;
      cur_mult *= cur_mult;
      delta >>= 1u;
    }
    state = acc_mult * state + acc_plus;
  }
  __device__ uint32_t next() {
    const uint64_t old = state;
    state = old * kPcgMult + inc;
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
  }
  __device__ double uniform() { return ((double)next() + 0.5) * (1.0 / 4294967296.0); }
};

__global__ void normals_kernel(uint64_t seed, uint64_t stream, long n, float *out) {
  const long npairs = (n + 1) / 2;
  const long p0 = ((long)blockIdx.x * blockDim.x + threadIdx.x) * kPairsPerThread;
  if (p0 >= npairs) return;
  DevPcg32 g;
  g.seed(seed, stream);
  g.advance(2u * (uint64_t)p0);
  const long p1 = p0 + kPairsPerThread < npairs ? p0 + kPairsPerThread : npairs;
  for (long p = p0; p < p1; ++p) {
    const double u1 = g.uniform(), u2 = g.uniform();
    const double r = sqrt(-2.0 * log(u1));
    const double th = 2.0 * 3.14159265358979323846 * u2;
    out[2 * p] = (float)(r * cos(th));
    if (2 * p + 1 < n) out[2 * p + 1] = (float)(r * sin(th));
  }
}

}  // namespace

hipError_t launch_normals(uint64_t seed, uint64_t stream, long n, float *out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const long npairs = (n + 1) / 2;
  const long threads = (npairs + kPairsPerThread - 1) / kPairsPerThread;
  const long blocks = (threads + 255) / 256;
  hipLaunchKernelGGL(normals_kernel, dim3((unsigned)blocks), dim3(256), 0, st, seed, stream, n, out);
  return hipGetLastError();
}

}  // namespace dasp
