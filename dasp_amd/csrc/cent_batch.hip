// cent_batch.hip - multi-WFS persistent kernels (batch_kernel.h) for the NP = 8 and NP = 16 shapes, and the shape dispatch.
#include "batch_kernel.h"

namespace dasp {

// (NP, fftsize, bin factor) of the batch kernels; NP = 32 shapes live in cent_batch32.hip.
#define DASP_BATCH_SMALL(X) X(8, 16, 2) X(16, 32, 2) X(16, 32, 4) X(8, 16, 1) X(16, 32, 1)
#define DASP_BATCH_LARGE(X) X(32, 64, 2) X(32, 64, 4) X(32, 64, 8)
// with the fused LGS spot convolution
#define DASP_BATCH_SPOT(X) X(8, 16, 2) X(16, 32, 2)

hipError_t launch_batch32(const BatchArgs &a, int F, int B, int waves, int chunk, int fast, long wgs, hipStream_t st);
int resident_batch32(int F, int B, int waves, int fast);

bool batch_supported(int np, int F, int nimg, bool spot) {
  if (nimg <= 0 || F % nimg) return false;
  const int b = F / nimg;
#define DASP_BATCH_HAS(A, Bf, Bb) \
  if (np == A && F == Bf && b == Bb) return true;
  if (spot) {
    DASP_BATCH_SPOT(DASP_BATCH_HAS)
  } else {
    DASP_BATCH_SMALL(DASP_BATCH_HAS)
    DASP_BATCH_LARGE(DASP_BATCH_HAS)
  }
#undef DASP_BATCH_HAS
  return false;
}

hipError_t launch_batch(const BatchArgs &a, int np, int F, int nimg, bool spot, int waves, int chunk, int fast,
                        long wgs, hipStream_t st) {
  if (a.total_groups <= 0) return hipSuccess;
  if (wgs <= 0 || wgs > 0x7fffffffL || !batch_supported(np, F, nimg, spot)) return hipErrorInvalidValue;
  const int b = F / nimg;
  if (np == 32) return launch_batch32(a, F, b, waves, chunk, fast, wgs, st);
#define DASP_BATCH_CASE(A, Bf, Bb) \
  if (np == A && F == Bf && b == Bb) return launch_batch_t<A, Bf, Bb, false>(a, waves, chunk, fast, wgs, st);
#define DASP_BATCH_SPOT_CASE(A, Bf, Bb) \
  if (np == A && F == Bf && b == Bb) return launch_batch_t<A, Bf, Bb, true>(a, waves, chunk, fast, wgs, st);
  if (spot) {
    DASP_BATCH_SPOT(DASP_BATCH_SPOT_CASE)
  } else {
    DASP_BATCH_SMALL(DASP_BATCH_CASE)
  }
#undef DASP_BATCH_CASE
#undef DASP_BATCH_SPOT_CASE
  return hipErrorInvalidValue;
}

int batch_resident_wgs(int np, int F, int nimg, bool spot, int waves, int fast) {
  if (!batch_supported(np, F, nimg, spot)) return 0;
  const int b = F / nimg;
  if (np == 32) return resident_batch32(F, b, waves, fast);
#define DASP_BATCH_OCC(A, Bf, Bb) \
  if (np == A && F == Bf && b == Bb) return resident_batch_t<A, Bf, Bb, false>(waves, fast);
#define DASP_BATCH_SPOT_OCC(A, Bf, Bb) \
  if (np == A && F == Bf && b == Bb) return resident_batch_t<A, Bf, Bb, true>(waves, fast);
  if (spot) {
    DASP_BATCH_SPOT(DASP_BATCH_SPOT_OCC)
  } else {
    DASP_BATCH_SMALL(DASP_BATCH_OCC)
  }
#undef DASP_BATCH_OCC
#undef DASP_BATCH_SPOT_OCC
  return 0;
}
}  // namespace dasp
