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

// Persistent waves: wave w handles item groups w, w + nwaves, ... and loads the next group's
// phase and pupil while it transforms the current one.
template <int NP, int F, int B, bool FAST, bool SPOT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NP <= 16 ? 2 : 1)))
void fused_cog_kernel(FusedArgs a) {
  using S = FusedShape<NP, F, B>;
  constexpr int SPW = S::SPW, SUB = stage_per_subap<NP, F, SPOT>();
  __shared__ cf stage[S::WAVES][SPW * SUB];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int sl = lane / NP;
  const int j = lane - sl * NP;   // pass 1: phase column; pass 2: row group m
  const long ngroups = (a.nitems + SPW - 1) / SPW;
  const long stride = (long)gridDim.x * S::WAVES;
  long g = (long)blockIdx.x * S::WAVES + w;
  if (g >= ngroups) return;       // wave-uniform
  cf *st = &stage[w][sl * SUB];
  PhaseCol<NP> cur;
  load_item<NP>(a, g * SPW + sl, j, cur);
  for (; g < ngroups; g += stride) {
    PhaseCol<NP> nxt;
    load_item<NP>(a, (g + stride) * SPW + sl, j, nxt);   // items past the end load zeros
    fused_item<NP, F, B, FAST, SPOT>(a, g * SPW + sl, j, w, sl, cur, st);
    cur = nxt;
  }
}

template <int NP, int F, int B, bool SPOT>
static hipError_t launch_fused_t(const FusedArgs &a, hipStream_t st) {
  using S = FusedShape<NP, F, B>;
  const long per_block = (long)S::WAVES * S::SPW;
  long blocks = (a.nitems + per_block - 1) / per_block;
  if (blocks <= 0) return hipSuccess;
  // persistent grid: enough workgroups for every CU to hold several, then grid-stride
  const long cap = 256L * 4;
  if (blocks > cap) blocks = cap;
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

hipError_t launch_fused(const FusedArgs &a, int np, int F, bool spot, hipStream_t st) {
  const int b = F / a.nimg;
#define DASP_FUSED_CASE(A, Bf, C) \
  if (np == A && F == Bf && b == C) return launch_fused_t<A, Bf, C, false>(a, st);
#define DASP_FUSED_SPOT_CASE(A, Bf, C) \
  if (np == A && F == Bf && b == C) return launch_fused_t<A, Bf, C, true>(a, st);
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
