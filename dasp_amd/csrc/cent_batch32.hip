// cent_batch32.hip - multi-WFS persistent kernels (batch_kernel.h) for the NP = 32 shapes (a translation unit of its own so the build compiles it in parallel).
#include "batch_kernel.h"

namespace dasp {

#define DASP_BATCH_LARGE(X) X(32, 64, 2) X(32, 64, 4) X(32, 64, 8)

hipError_t launch_batch32(const BatchArgs &a, int F, int b, int waves, int chunk, int fast, long wgs, hipStream_t st) {
#define DASP_BATCH_CASE(A, Bf, Bb) \
  if (F == Bf && b == Bb) return launch_batch_t<A, Bf, Bb, false>(a, waves, chunk, fast, wgs, st);
  DASP_BATCH_LARGE(DASP_BATCH_CASE)
#undef DASP_BATCH_CASE
  return hipErrorInvalidValue;
}

int resident_batch32(int F, int b, int waves, int fast) {
#define DASP_BATCH_OCC(A, Bf, Bb) \
  if (F == Bf && b == Bb) return resident_batch_t<A, Bf, Bb, false>(waves, fast);
  DASP_BATCH_LARGE(DASP_BATCH_OCC)
#undef DASP_BATCH_OCC
  return 0;
}

}  // namespace dasp
