// recon_internal.h - arguments of the reconstructor kernels (recon_mvm.hip) and their launcher.
#pragma once
#include <hip/hip_runtime.h>

namespace dasp {

constexpr int kReconRows = 256;    // slope rows per block of the layout-0 product (4 waves x 64)
constexpr int kReconMaxF = 8;      // frames per launch (the matrix is read once per launch)

struct ReconArgs {
  const float *mx;       // reconmx, borrowed device memory
  long ld;               // row stride (floats)
  int ncents, nacts;
  const int *index;      // (ncents) slope index within one frame of the slope buffer
  const float *slopes;   // frame f at slopes + f * slope_stride
  long slope_stride;
  float *acts;           // frame f at acts + f * acts_stride, updated in place
  long acts_stride;
  const float *gains;    // (nacts) or nullptr: gain_s for every actuator
  float gain_s;
  const float *decay;    // (nacts) or nullptr: decay_s for every actuator
  float decay_s;
  float *gathered;       // scratch: the gathered slope vectors (frames x gstride)
  long gstride;          // ncents rounded up to a multiple of 4
  float *part;           // scratch: partial sums of the layout-0 product (chunks x frames x nacts)
};

hipError_t launch_recon(const ReconArgs &a, int layout, int F, hipStream_t st);

}  // namespace dasp
