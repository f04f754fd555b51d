// recon_api.cpp - C ABI of the device reconstructor step (include/dasp_cent.h, dasp_recon_*).
// The product of science/tomoRecon.py's dense reconstructors (:1131-1174) over slopes the engine
// left in HBM; see recon_mvm.hip for the kernels.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/dasp_cent.h"
#include "recon_internal.h"

using namespace dasp;

namespace {
thread_local std::string r_err;
int rfail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  r_err = buf;
  return code;
}
}  // namespace

struct dasp_recon {
  int ncents = 0, nacts = 0, layout = 0;
  int maxindex = 0;
  const float *mx = nullptr;
  long ld = 0;
  int *d_index = nullptr;
  float *d_gains = nullptr, *d_decay = nullptr, *d_part = nullptr;
  float gain = 1.f, decay = 1.f;
};

#define RCHK(expr)                                                                              \
  do {                                                                                          \
    hipError_t _e = (expr);                                                                     \
    if (_e != hipSuccess) return rfail(DASP_CENT_EHIP, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

extern "C" {

const char *dasp_recon_last_error(void) { return r_err.c_str(); }

int dasp_recon_free(dasp_recon_t *r) {
  if (!r) return 0;
  if (r->d_index) (void)hipFree(r->d_index);
  if (r->d_gains) (void)hipFree(r->d_gains);
  if (r->d_decay) (void)hipFree(r->d_decay);
  if (r->d_part) (void)hipFree(r->d_part);
  delete r;
  return 0;
}

int dasp_recon_create(int ncents, int nacts, int layout, const float *d_reconmx, long ld, const int *h_index,
                      dasp_recon_t **out) {
  if (!out) return rfail(DASP_CENT_EINVAL, "null output pointer");
  *out = nullptr;
  if (ncents < 1 || nacts < 1 || !d_reconmx || !h_index || (layout != 0 && layout != 1))
    return rfail(DASP_CENT_EINVAL, "recon: need ncents, nacts >= 1, layout 0 or 1, a matrix and an index");
  if (ld < (layout == 0 ? nacts : ncents)) return rfail(DASP_CENT_EINVAL, "recon: row stride %ld too small", ld);
  int maxindex = 0;
  for (int c = 0; c < ncents; ++c) {
    if (h_index[c] < 0) return rfail(DASP_CENT_EINVAL, "recon: negative slope index at %d", c);
    if (h_index[c] > maxindex) maxindex = h_index[c];
  }
  dasp_recon *r = new dasp_recon;
  r->maxindex = maxindex;
  r->ncents = ncents;
  r->nacts = nacts;
  r->layout = layout;
  r->mx = d_reconmx;
  r->ld = ld;
  const long nchunks = (ncents + kReconRows - 1) / kReconRows;
  const long gstride = (ncents + 3) & ~3L;
  const long part = (long)kReconMaxF * gstride + (layout == 0 ? nchunks * kReconMaxF * nacts : 0);
  hipError_t e = hipMalloc((void **)&r->d_index, sizeof(int) * ncents);
  if (e == hipSuccess) e = hipMemcpy(r->d_index, h_index, sizeof(int) * ncents, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc((void **)&r->d_part, sizeof(float) * part);
  if (e != hipSuccess) {
    dasp_recon_free(r);
    return rfail(DASP_CENT_EHIP, "recon set-up: %s", hipGetErrorString(e));
  }
  *out = r;
  return 0;
}

int dasp_recon_set_gains(dasp_recon_t *r, const float *h_gains, float gain, const float *h_decay, float decay) {
  if (!r) return rfail(DASP_CENT_EINVAL, "null recon");
  r->gain = gain;
  r->decay = decay;
  for (int k = 0; k < 2; ++k) {
    const float *src = k == 0 ? h_gains : h_decay;
    float *&dst = k == 0 ? r->d_gains : r->d_decay;
    if (!src) {
      if (dst) (void)hipFree(dst);
      dst = nullptr;
      continue;
    }
    if (!dst) RCHK(hipMalloc((void **)&dst, sizeof(float) * r->nacts));
    RCHK(hipMemcpy(dst, src, sizeof(float) * r->nacts, hipMemcpyHostToDevice));
  }
  return 0;
}

int dasp_recon_run(dasp_recon_t *r, const float *d_slopes, long slope_stride, int nframes, float *d_acts,
                   long acts_stride, void *stream) {
  if (!r) return rfail(DASP_CENT_EINVAL, "null recon");
  if (!d_slopes || !d_acts || nframes < 1) return rfail(DASP_CENT_EINVAL, "recon: slopes, acts and nframes >= 1 required");
  // every gathered entry lies inside one frame of the slope buffer, every update inside one
  // frame of the actuator buffer
  if (slope_stride <= r->maxindex)
    return rfail(DASP_CENT_EINVAL, "recon: slope_stride %ld must exceed the largest slope index %d", slope_stride, r->maxindex);
  if (nframes > 1 && acts_stride < r->nacts) return rfail(DASP_CENT_EINVAL, "recon: acts_stride too small");
  ReconArgs a{};
  a.mx = r->mx;
  a.ld = r->ld;
  a.ncents = r->ncents;
  a.nacts = r->nacts;
  a.index = r->d_index;
  a.gains = r->d_gains;
  a.gain_s = r->gain;
  a.decay = r->d_decay;
  a.decay_s = r->decay;
  a.gstride = (r->ncents + 3) & ~3L;
  a.gathered = r->d_part;
  a.part = r->d_part + (long)kReconMaxF * a.gstride;
  a.slope_stride = slope_stride;
  a.acts_stride = acts_stride;
  for (int f0 = 0; f0 < nframes; f0 += kReconMaxF) {   // the matrix is read once per kReconMaxF frames
    const int F = nframes - f0 < kReconMaxF ? nframes - f0 : kReconMaxF;
    a.slopes = d_slopes + (long)f0 * slope_stride;
    a.acts = d_acts + (long)f0 * acts_stride;
    RCHK(launch_recon(a, r->layout, F, (hipStream_t)stream));
  }
  return 0;
}

}  // extern "C"
