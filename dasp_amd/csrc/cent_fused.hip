// cent_fused.hip - the fused FP32 path: phase -> |FFT|^2 -> binned detector image -> noise ->
// threshold -> CoG, one pass per subaperture with nothing but the slopes (and optionally the
// images) written back to HBM.
//
// Covers the configurations where clipsize == fftsize, the binning is an integer block sum,
// and the centroid is a (weighted) CoG: computeHll, addPowerSpecCentral, binImage, the
// normalisation, sky, poissonise/readCCD, pixel power and computeCoG of cmod/src/centmodule.c
// (:430-509, :609-651, :1829-1897, :873-971), and the LGS spot convolution (doConvolution
// :524-564) when psfsize == clipsize == fftsize. Everything else takes the general path
// (cent_form.hip + cent_backend.hip).

#include "fused_item.h"

namespace dasp {

// Persistent waves over groups of SPW consecutive active subapertures of one frame. XCD-aware:
// workgroup b runs on XCD b % 8 (round-robin dispatch), and XCD x owns the x-th eighth of the
// subaperture groups for every frame, so each XCD's L2 keeps its slice of the pupil (and of any
// per-subaperture table) across frames instead of every frame re-reading the whole pupil. Inside
// an XCD the groups are taken with a grid stride, frame-major, or for n = 32 (4 KB of pupil per
// sub-aperture, where the slice's pupil competes with the streamed phase for L2)
// sub-aperture-group-major: every frame of a group, then the next group, so that the XCD's waves
// work on a few groups at a time (1 % faster at n32/nfft64, 1 % slower at n16/nfft32 with noise,
// same-box A/B). With NP = 32 or a spot kernel (two waves per SIMD or fewer either way) each wave
// loads its next group's phase and pupil while it transforms the current one; the NP <= 16 shapes
// without a spot kernel load each group when they start it and run three waves per SIMD instead
// (the prefetched column held 30-40 VGPRs more and kept them at two): the other waves hide the
// load latency, 0.209 -> 0.187 ms per 80x80 nfft32 step and 0.168 -> 0.141 ms per 40x40 noisy
// step (same-box A/B).
constexpr int kXcds = 8;

template <int NP, bool SPOT>
constexpr bool kPrefetch = SPOT || NP >= 32;

template <int NP, int F, int B, bool FAST, bool SPOT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NP <= 16 ? (kPrefetch<NP, SPOT> ? 2 : 3) : 1)))
void fused_cog_kernel(FusedArgs a_) {
  // Read the arguments through a constant-address-space pointer to the kernel argument segment:
  // the compiler then reloads fields where they are used instead of holding them in SGPRs
  // (half the SGPR spills, 4-5 % faster at n16/nfft32, measured).
  typedef const __attribute__((address_space(4))) FusedArgs KernArgs;
  const FusedArgs &a = *(const FusedArgs *)(KernArgs *)__builtin_amdgcn_kernarg_segment_ptr();
  using S = FusedShape<NP, F, B>;
  constexpr int SPW = S::SPW;
  __shared__ cf stage[S::WAVES][wave_lds_cf<NP, F, B, SPOT>()];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int sl = lane / NP;
  const int j = lane - sl * NP;   // pass 1: phase column; pass 2: row group m
  const int xcd = blockIdx.x % kXcds;
  const long nsg = (a.nactive + SPW - 1) / SPW;              // subaperture groups per frame
  const long sg0 = nsg * xcd / kXcds, nsgx = nsg * (xcd + 1) / kXcds - sg0;
  const long nframes = a.nitems / a.nactive;
  const long ngx = nsgx * nframes;                           // this XCD's groups
  const long stride = (long)(gridDim.x / kXcds) * S::WAVES;
  long q = (long)(blockIdx.x / kXcds) * S::WAVES + w;
  zero_inactive_slopes(a);
  if (q >= ngx) return;           // wave-uniform
  // group q = g * nframes + f (sub-aperture-group-major) or q = f * nsgx + g (frame-major) for
  // group sg0 + g of frame f; q advances by a fixed stride, so frame and group advance by fixed
  // steps: one scalar division per wave
  constexpr bool kGroupMajor = NP >= 32;
  const long inner = kGroupMajor ? nframes : nsgx;   // the faster-moving index's range
  const int ninner = (int)inner;
  const int qo = (int)(q / inner), qi = (int)(q - (long)qo * inner);
  int gcur = kGroupMajor ? qo : qi, fcur = kGroupMajor ? qi : qo;
  const int stepO = (int)(stride / inner), stepI = (int)(stride - (long)stepO * inner);
  auto item_of = [&](long qq, int f, int g) -> ItemRef {   // wave-uniform arguments
    return item_at(a, qq < ngx ? f : -1, (int)((sg0 + g) * SPW) + sl);
  };
  cf *wl = &stage[w][0];
  if constexpr (!kPrefetch<NP, SPOT>) {
    for (; q < ngx; q += stride) {
      const ItemRef it = item_of(q, fcur, gcur);
      PhaseCol<NP> cur;
      load_item<NP, true>(a, it, j, cur, fcur);
      fused_item<NP, F, B, FAST, SPOT>(a, it, j, w, sl, cur, wl);
      int ni = (kGroupMajor ? fcur : gcur) + stepI, no = (kGroupMajor ? gcur : fcur) + stepO;
      if (ni >= ninner) {
        ni -= ninner;
        ++no;
      }
      fcur = kGroupMajor ? ni : no;
      gcur = kGroupMajor ? no : ni;
    }
  } else {
    PhaseCol<NP> cur;
    ItemRef it = item_of(q, fcur, gcur);
    load_item<NP, true>(a, it, j, cur, fcur);
    for (; q < ngx; q += stride) {
      int ni = (kGroupMajor ? fcur : gcur) + stepI, no = (kGroupMajor ? gcur : fcur) + stepO;
      if (ni >= ninner) {
        ni -= ninner;
        ++no;
      }
      const int nf = kGroupMajor ? ni : no, ng = kGroupMajor ? no : ni;
      PhaseCol<NP> nxt;
      const bool more = q + stride < ngx;
      const ItemRef itn = item_of(q + stride, nf, ng);
      load_item<NP, true>(a, itn, j, nxt, more ? nf : 0);   // past the end: item 0, unused
      fused_item<NP, F, B, FAST, SPOT>(a, it, j, w, sl, cur, wl);
      cur = nxt;
      it = itn;
      fcur = nf;
      gcur = ng;
    }
  }
}

template <int NP, int F, int B, bool SPOT>
static hipError_t launch_fused_t(const FusedArgs &a, hipStream_t st, long grid) {
  using S = FusedShape<NP, F, B>;
  if (a.nitems <= 0 || a.nactive <= 0) return hipSuccess;
  const long nsg = (a.nactive + S::SPW - 1) / S::SPW;
  const long groups = nsg * (a.nitems / a.nactive);
  // persistent grid: enough workgroups for every CU to hold several (or the tuned grid)
  const long blocks = launch_grid((groups + S::WAVES - 1) / S::WAVES, 256L * 4, grid);
  if (a.fast_trig)
    hipLaunchKernelGGL((fused_cog_kernel<NP, F, B, true, SPOT>), dim3((unsigned)blocks), dim3(64 * S::WAVES), 0, st, a);
  else
    hipLaunchKernelGGL((fused_cog_kernel<NP, F, B, false, SPOT>), dim3((unsigned)blocks), dim3(64 * S::WAVES), 0, st, a);
  return hipGetLastError();
}

#define DASP_FUSED_SHAPES(X) X(8, 16, 2) X(16, 32, 2) X(32, 64, 2) X(32, 64, 4) X(32, 64, 8) X(16, 32, 4) X(8, 16, 1) X(16, 32, 1)
// shapes with the fused LGS spot convolution (psfsize == clipsize == fftsize)
#define DASP_FUSED_SPOT_SHAPES(X) X(8, 16, 2) X(16, 32, 2)

bool fused_supported(int np, int F, int nimg, bool spot) {
  if (nimg <= 0 || F % nimg) return false;
  const int b = F / nimg;
#define DASP_FUSED_HAS(A, Bf, C) \
  if (np == A && F == Bf && b == C) return true;
  if (spot) {
    DASP_FUSED_SPOT_SHAPES(DASP_FUSED_HAS)
  } else {
    DASP_FUSED_SHAPES(DASP_FUSED_HAS)
  }
#undef DASP_FUSED_HAS
  return false;
}

hipError_t launch_fused(const FusedArgs &a, int np, int F, bool spot, hipStream_t st, long grid) {
  // the argument block comes from another translation unit: refuse, rather than divide by, a
  // detector size that cannot be right
  if (a.nimg <= 0 || F % a.nimg != 0) return hipErrorInvalidValue;
  const int b = F / a.nimg;
#define DASP_FUSED_CASE(A, Bf, C) \
  if (np == A && F == Bf && b == C) return launch_fused_t<A, Bf, C, false>(a, st, grid);
#define DASP_FUSED_SPOT_CASE(A, Bf, C) \
  if (np == A && F == Bf && b == C) return launch_fused_t<A, Bf, C, true>(a, st, grid);
  if (spot) {
    DASP_FUSED_SPOT_SHAPES(DASP_FUSED_SPOT_CASE)
  } else {
    DASP_FUSED_SHAPES(DASP_FUSED_CASE)
  }
#undef DASP_FUSED_CASE
#undef DASP_FUSED_SPOT_CASE
  return hipErrorInvalidValue;
}

}  // namespace dasp
