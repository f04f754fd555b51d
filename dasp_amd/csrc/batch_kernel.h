// batch_kernel.h - the multi-WFS persistent kernel, one instantiation per fused shape.
//
// The host lays the WFS of one shape out as a queue of groups (64/NP work items each, a WFS's
// groups contiguous; items frame-major inside a WFS as in the single-WFS kernel, or
// subaperture-major for spot kernels). Waves walk the queue in increasing order, statically or
// by atomic claims (WaveQueue); the next group's phase and pupil are loaded while the current
// group is transformed, also across a WFS boundary. With claims, the last wave to finish resets
// the counters, so a captured step graph replays the launch without a memset node.
//
// One kernel per shape rather than one for all shapes: a kernel's register allocation covers
// every shape compiled into it, and a five-shape kernel spills where each single-shape kernel
// fits (NP <= 16: two waves per SIMD; NP = 32: one).
#pragma once
#include "fused_item.h"

namespace dasp {

// Work schedule of one wave over the queue. chunk == 0: static, wave w of W takes entries
// w, w + W, w + 2W, ... (no atomics; balanced when entries cost the same, as they do within one
// shape). chunk > 0: dynamic, the wave claims chunk consecutive entries per atomic on the
// queue counter (balances uneven entries; one atomic per chunk bounds the contention on the
// counter). Called from whole-wave code, so lane 0 is the first active lane.
struct WaveQueue {
  unsigned long long *ctr;
  long stride;   // static: total waves
  int chunk;
  long lim = 0;  // dynamic: end of the current claim
  __device__ long first(long wave) {
    if (chunk == 0) return wave;
    return claim();
  }
  __device__ long next(long g) {
    if (chunk == 0) return g + stride;
    if (g + 1 < lim) return g + 1;
    return claim();
  }
  __device__ long claim() {
    unsigned long long g = 0;
    if ((threadIdx.x & 63) == 0) g = atomicAdd(ctr, (unsigned long long)chunk);
    const long start = wave_uniform(g);
    lim = start + chunk;
    return start;
  }
};

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

// Queue order inside a WFS: frame-major, or subaperture-major for WFS with spot kernels (a
// group's items then share one subaperture's spot spectrum).
template <bool SPOT>
__device__ __forceinline__ ItemRef flat_item(const FusedArgs &a, long first, int sl) {
  if constexpr (SPOT) return item_flat_by_subap(a, first, sl);
  else return item_flat(a, first, sl);
}

// As in the single-WFS kernel: the NP <= 16 shapes without a spot kernel load each group when the
// wave starts it and run three waves per SIMD; the spot shapes (two waves by their LDS) and NP = 32
// load the next group while they transform the current one.
template <int NP, bool SPOT>
constexpr bool kBatchPrefetch = SPOT || NP >= 32;

template <int NP, int F, int B, int WAVES, bool FAST, bool SPOT>
__global__ __launch_bounds__(64 * WAVES)
__attribute__((amdgpu_waves_per_eu(NP <= 16 ? (kBatchPrefetch<NP, SPOT> ? 2 : 3) : 1)))
void batch_kernel(BatchArgs b, int chunk) {
  constexpr int SPW = 64 / NP;
  __shared__ cf stage[WAVES][wave_lds_cf<NP, F, B, SPOT>()];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int sl = lane / NP;
  const int j = lane - sl * NP;
  cf *wl = &stage[w][0];
  WaveQueue q{b.counters, (long)gridDim.x * WAVES, chunk};
  long g = q.first((long)blockIdx.x * WAVES + w);
  if (g < b.total_groups && !kBatchPrefetch<NP, SPOT>) {
    int d = desc_of(b, g);
    long end = desc_end(b, d);
    while (true) {
      PhaseCol<NP> cur;
      const ItemRef it = flat_item<SPOT>(desc_args(b, d), (g - b.desc[d].group0) * SPW, sl);
      load_item<NP>(desc_args(b, d), it, j, cur);
      fused_item<NP, F, B, FAST, SPOT>(desc_args(b, d), it, j, w, sl, cur, wl);
      const long gn = q.next(g);   // entries only grow: gn > g
      if (gn >= b.total_groups) break;
      if (gn >= end) {
        d = desc_of(b, gn);
        end = desc_end(b, d);
      }
      g = gn;
    }
  } else if (g < b.total_groups) {
    int d = desc_of(b, g);
    long base = b.desc[d].group0, end = desc_end(b, d);
    PhaseCol<NP> cur;
    ItemRef it = flat_item<SPOT>(desc_args(b, d), (g - base) * SPW, sl);
    load_item<NP>(desc_args(b, d), it, j, cur);
    while (true) {
      const long gn = q.next(g);   // entries only grow: gn > g
      const bool more = gn < b.total_groups;
      const int dn = more && gn >= end ? desc_of(b, gn) : d;
      PhaseCol<NP> nxt;
      ItemRef itn{0, 0, 0, false};
      if (more) {
        itn = flat_item<SPOT>(desc_args(b, dn), (gn - b.desc[dn].group0) * SPW, sl);
        load_item<NP>(desc_args(b, dn), itn, j, nxt);
      }
      fused_item<NP, F, B, FAST, SPOT>(desc_args(b, d), it, j, w, sl, cur, wl);
      if (!more) break;
      cur = nxt;
      it = itn;
      g = gn;
      if (dn != d) {
        d = dn;
        base = b.desc[d].group0;
        end = desc_end(b, d);
      }
    }
  }
  // dynamic schedule: every wave reaches this point once; the last one resets the counters
  if (chunk > 0 && lane == 0) {
    __threadfence();
    const unsigned long long nw = (unsigned long long)gridDim.x * WAVES;
    if (atomicAdd(b.counters + 1, 1ull) == nw - 1) {
      atomicExch(b.counters, 0ull);
      atomicExch(b.counters + 1, 0ull);
    }
  }
}

// Launcher and occupancy query for one shape: waves per workgroup 1, 2 or 4.
template <int NP, int F, int B, bool SPOT>
hipError_t launch_batch_t(const BatchArgs &a, int waves, int chunk, int fast, long wgs, hipStream_t st) {
#define DASP_BATCH_LAUNCH(W)                                                                           \
  if (waves == W) {                                                                                   \
    if (fast) hipLaunchKernelGGL((batch_kernel<NP, F, B, W, true, SPOT>), dim3((unsigned)wgs), dim3(64 * W), 0, st, a, chunk); \
    else hipLaunchKernelGGL((batch_kernel<NP, F, B, W, false, SPOT>), dim3((unsigned)wgs), dim3(64 * W), 0, st, a, chunk);     \
    return hipGetLastError();                                                                         \
  }
  DASP_BATCH_LAUNCH(1)
  DASP_BATCH_LAUNCH(2)
  DASP_BATCH_LAUNCH(4)
#undef DASP_BATCH_LAUNCH
  return hipErrorInvalidValue;
}

template <int NP, int F, int B, bool SPOT>
int resident_batch_t(int waves, int fast) {
  int nb = 0;
  hipError_t e = hipErrorInvalidValue;
#define DASP_BATCH_OCC(W)                                                                                     \
  if (waves == W)                                                                                            \
    e = fast ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<NP, F, B, W, true, SPOT>, 64 * W, 0)   \
             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<NP, F, B, W, false, SPOT>, 64 * W, 0);
  DASP_BATCH_OCC(1)
  DASP_BATCH_OCC(2)
  DASP_BATCH_OCC(4)
#undef DASP_BATCH_OCC
  return e == hipSuccess ? nb : 0;
}

}  // namespace dasp
