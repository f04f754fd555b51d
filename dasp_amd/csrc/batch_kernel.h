// batch_kernel.h - the multi-WFS persistent kernel, one instantiation per fused shape.
//
// The host lays the WFS of one shape out as a queue of groups (64/NP work items each, a WFS's
// groups contiguous, items frame-major inside a WFS as in the single-WFS kernel). Waves take
// groups from a device counter in increasing order, so a WFS boundary is crossed once per wave.
// The next group's phase and pupil are loaded while the current group is transformed, also
// across a WFS boundary. The last wave to finish resets the counters, so a captured step graph
// replays the launch without a memset node.
//
// One kernel per shape rather than one for all shapes: a kernel's register allocation covers
// every shape compiled into it, and a five-shape kernel spills where each single-shape kernel
// fits (NP <= 16: two waves per SIMD; NP = 32: one).
#pragma once
#include "fused_item.h"

namespace dasp {

__device__ __forceinline__ long wave_uniform(unsigned long long v) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (long)(((unsigned long long)hi << 32) | lo);
}

// Next queue entry. Called from whole-wave code, so lane 0 is the first active lane.
__device__ __forceinline__ long grab_group(unsigned long long *ctr, int lane) {
  unsigned long long g = 0;
  if (lane == 0) g = atomicAdd(ctr, 1ull);
  return wave_uniform(g);
}

// Descriptor holding queue entry g: the last one whose first group is <= g.
__device__ __forceinline__ int desc_of(const BatchArgs &b, long g) {
  int lo = 0, hi = b.ndesc - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (b.desc[mid].group0 <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Arguments of descriptor d, read through the constant address space: the loads are then
// invariant scalar loads the compiler re-issues where needed instead of holding ~70 SGPRs.
__device__ __forceinline__ const FusedArgs &desc_args(const BatchArgs &b, int d) {
  typedef const __attribute__((address_space(4))) BatchDesc ConstDesc;
  return *(const FusedArgs *)&((ConstDesc *)b.desc)[d].args;
}

__device__ __forceinline__ long desc_end(const BatchArgs &b, int d) {
  return d + 1 < b.ndesc ? b.desc[d + 1].group0 : b.total_groups;
}

template <int NP, int F, int B, int WAVES, bool FAST>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(NP <= 16 ? 2 : 1)))
void batch_kernel(BatchArgs b) {
  constexpr int SPW = 64 / NP, TS = NP + 1;
  __shared__ cf stage[WAVES][SPW * NP * TS];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int sl = lane / NP;
  const int j = lane - sl * NP;
  cf *st = &stage[w][sl * NP * TS];
  long g = grab_group(b.counters, lane);
  if (g < b.total_groups) {
    int d = desc_of(b, g);
    long base = b.desc[d].group0, end = desc_end(b, d);
    PhaseCol<NP> cur;
    load_item<NP>(desc_args(b, d), (g - base) * SPW + sl, j, cur);
    while (true) {
      const long gn = grab_group(b.counters, lane);   // entries only grow: gn > g
      const bool more = gn < b.total_groups;
      const int dn = more && gn >= end ? desc_of(b, gn) : d;
      PhaseCol<NP> nxt;
      if (more) load_item<NP>(desc_args(b, dn), (gn - b.desc[dn].group0) * SPW + sl, j, nxt);
      fused_item<NP, F, B, FAST>(desc_args(b, d), (g - base) * SPW + sl, j, w, sl, cur, st);
      if (!more) break;
      cur = nxt;
      g = gn;
      if (dn != d) {
        d = dn;
        base = b.desc[d].group0;
        end = desc_end(b, d);
      }
    }
  }
  // every wave reaches this point once; the last one resets the queue for the next launch
  if (lane == 0) {
    __threadfence();
    const unsigned long long nw = (unsigned long long)gridDim.x * WAVES;
    if (atomicAdd(b.counters + 1, 1ull) == nw - 1) {
      atomicExch(b.counters, 0ull);
      atomicExch(b.counters + 1, 0ull);
    }
  }
}

// Launcher and occupancy query for one shape: waves per workgroup 1, 2 or 4.
template <int NP, int F, int B>
hipError_t launch_batch_t(const BatchArgs &a, int waves, int fast, long wgs, hipStream_t st) {
#define DASP_BATCH_LAUNCH(W)                                                                           \
  if (waves == W) {                                                                                   \
    if (fast) hipLaunchKernelGGL((batch_kernel<NP, F, B, W, true>), dim3((unsigned)wgs), dim3(64 * W), 0, st, a); \
    else hipLaunchKernelGGL((batch_kernel<NP, F, B, W, false>), dim3((unsigned)wgs), dim3(64 * W), 0, st, a);     \
    return hipGetLastError();                                                                         \
  }
  DASP_BATCH_LAUNCH(1)
  DASP_BATCH_LAUNCH(2)
  DASP_BATCH_LAUNCH(4)
#undef DASP_BATCH_LAUNCH
  return hipErrorInvalidValue;
}

template <int NP, int F, int B>
int resident_batch_t(int waves, int fast) {
  int nb = 0;
  hipError_t e = hipErrorInvalidValue;
#define DASP_BATCH_OCC(W)                                                                                     \
  if (waves == W)                                                                                            \
    e = fast ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<NP, F, B, W, true>, 64 * W, 0)   \
             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<NP, F, B, W, false>, 64 * W, 0);
  DASP_BATCH_OCC(1)
  DASP_BATCH_OCC(2)
  DASP_BATCH_OCC(4)
#undef DASP_BATCH_OCC
  return e == hipSuccess ? nb : 0;
}

}  // namespace dasp
