// cent_api.cpp - host side of the gfx950 centroiding engine behind include/dasp_cent.h.
//
// Mirrors the state machine of cmod/src/centmodule.c (py_initialise :2858, py_run :3188,
// py_runSlope :3230, py_update :2573, py_free :3163) with the per-subaperture work moved to HIP
// kernels. Differences from the reference, all deliberate:
//  - no host thread pool; nthreads is accepted and ignored;
//  - noise comes from pre-drawn N(0,1) arrays (host PCG32/Box-Muller streams per readout) and
//    the inverse transform in noise.h, instead of in-module PCG32/GSL draws;
//  - arrays given to initialise/update are copied to HBM at that call (the phase and the input
//    image are re-read from the borrowed buffer at every run).
#include <hip/hip_runtime.h>
#include <sys/time.h>

#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <functional>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/dasp_cent.h"
#include "cent_internal.h"
#include "host_math.h"

using namespace dasp;

// Layout pins of the ABI structures (ctypes in dasp_amd/cmod/cent.py and the extension module
// mirror them); a change must bump DASP_CENT_ABI_VERSION and these numbers together.
static_assert(DASP_CENT_ABI_VERSION == 2, "ABI version and layout pins out of step");
static_assert(sizeof(dasp_cent_array) == 48, "dasp_cent_array layout changed");
static_assert(sizeof(dasp_cent_params) == 856, "dasp_cent_params layout changed");
static_assert(offsetof(dasp_cent_params, seed) == 168 && offsetof(dasp_cent_params, phs) == 176 &&
              offsetof(dasp_cent_params, subapLocation) == 808, "dasp_cent_params field offsets changed");
static_assert(sizeof(dasp_cent_batch_item) == 72, "dasp_cent_batch_item layout changed");
static_assert(DASP_CENT_NOALLOC_SCRATCH_BYTES == (long)kNoAllocScratchBytes, "fixed scratch size out of step");

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess) return fail(DASP_CENT_EHIP, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

// Set while dasp_cent_run_device_noalloc runs on this thread: a device buffer that would have to
// grow fails instead of calling hipFree/hipMalloc.
thread_local bool t_forbid_alloc = false;
thread_local bool t_alloc_refused = false;

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t nb) {
    if (nb <= bytes && p) return hipSuccess;
    if (t_forbid_alloc && nb > 0) {
      t_alloc_refused = true;
      return hipErrorOutOfMemory;
    }
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (nb == 0) return hipSuccess;
    hipError_t e = hipMalloc(&p, nb);
    if (e == hipSuccess) bytes = nb;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T *as() const { return reinterpret_cast<T *>(p); }
};

hipError_t upload(DevBuf &b, const void *src, size_t nb) {
  hipError_t e = b.ensure(nb);
  if (e != hipSuccess || nb == 0) return e;
  return hipMemcpy(b.p, src, nb, hipMemcpyHostToDevice);
}

long numel(const dasp_cent_array &a) {
  long n = 1;
  for (int i = 0; i < a.ndim; ++i) n *= a.dims[i];
  return n;
}

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

int pad_pow2(int v) {
  int r = 8;
  while (r < v) r <<= 1;
  return r;
}

}  // namespace

struct dasp_cent {
  dasp_cent_params p{};
  int nsubx = 0, nsubaps = 0, n = 0, F = 0, clip = 0, nimg = 0, ncen = 0;
  int psf = 0, padded = 0, pbf = 1, preBinOrig = 1, spotDim = 0;
  int binB = 0;   // integer binning factor when the binning operator is the plain block sum
  int corrsize = 0, corrPatternSize = 0, npup = 0;
  int atmosPhaseType = 0, interpolate = 0, inputPhaseSize = 0;
  int curInteg = 0, form = DASP_CENT_FORM_FFT32;
  int fastTrig = 1;   // reduced-argument hardware sin/cos (max abs error 3.8e-7, measured)
  unsigned long gen = 0;   // bumped by every update/set_* call: step graphs built on it are stale
  unsigned long seed = 1;
  unsigned long nreadouts = 0;
  float sig = 0.f, skybrightness = 0.f, readnoise = 0.f, readbg = 0.f, noiseFloor = 0.f;
  float pxlPower = 1.f, corrThresh = 0.f;
  int addPoisson = 0, calsource = 0, opticalBinning = 0, threshType = 0, imageOnly = 0;
  int correlationCentroiding = 0, useBrightest = 0, parabolicFit = 0, gaussianFit = 0;
  float gaussianMinVal = 0.f, gaussianReplaceVal = 0.f;
  int nintegrations = 1, nlatency = 0;
  int centWeightDim = 0, centWeightSize = 0;
  int calNSteps = 0, calNCoeff = 0;
  bool haveSigArr = false, haveSkyArr = false, haveRef = false, haveCalData = false;
  bool haveCalCoeff = false, haveUBArr = false, haveCorr = false, haveCentWeight = false;
  bool tablesDirty = true;
  bool allScalePos = false;   // every per-subaperture normalisation scale > 0 (upload_tables)
  bool relFloor = false;   // engine option: threshType 0 floor taken relative to the image maximum
  unsigned long scratchEpoch = 0;   // bumped whenever a scratch buffer of the handle moves
  void *noallocScratch = nullptr;   // the fixed scratch array on this handle's device
  std::vector<int> active, inactive, subflag;
  std::vector<float> pupfn, frac, sigArr, tilt, totsig, spot;
  int fitMxSize = -999;
  float fitMx[30]{};
  // borrowed host outputs / inputs
  float *h_phs = nullptr, *h_cents = nullptr, *h_bimg = nullptr, *h_corrimg = nullptr;
  const float *h_input = nullptr;
  float *h_outphase = nullptr;   // OUTPUTPHASEARR: receives the interpolated phase of active windows
  int inputDims = 0, inputNpxlx = 0;
  long phsElems = 0;
  // device state
  DevBuf d_phs, d_pup, d_tilt, d_active, d_winoff, d_inactive, d_scale, d_totsig, d_sigv, d_skyadd, d_skyarr;
  DevBuf d_poisson, d_binA, d_otf, d_corrK, d_corrKp, d_corrPat, d_centWeight, d_refCents, d_calData;
  DevBuf d_calBounds, d_calSteps, d_calCoeff, d_ubArr, d_input, d_subapLoc, d_cents, d_bimg;
  DevBuf d_corrimg, d_hll, d_zpois, d_zread, d_bscratch, d_cscratch;
  DevBuf d_otfk;   // spot kernel table of the fused kernels (psfsize == fftsize)
  DevBuf d_interp; // resampled phase grids when the input grid is not nsubx*phasesize
  std::vector<float> h_stage;
  // kernel timing
  bool timing = false;
  long gridImage = 0, gridCorr = 0;   // launch grids of the image and correlation kernels (0: default)
  std::vector<hipEvent_t> evs;   // start/stop pairs not yet harvested
  // side stream of the correlation stage, which runs frame chunks beside the image kernel's next
  // chunk (created on first use, default priority), and its fork / join events
  hipStream_t corrSide = nullptr;
  hipEvent_t corrFork = nullptr, corrJoin = nullptr;
  double kernel_ms = 0.0;
  long kernel_launches = 0;
};

// Grows a scratch buffer of the handle. When the buffer moves (a larger allocation replaces it),
// the handle's scratch epoch is bumped: a multi-WFS step graph that captured the old address is
// stale and is rebuilt at its next run.
static hipError_t grow_scratch(dasp_cent *c, DevBuf &b, size_t nb) {
  const void *old = b.p;
  const hipError_t e = b.ensure(nb);
  if (b.p != old) c->scratchEpoch++;
  return e;
}

static int harvest_events(dasp_cent *c) {
  for (size_t i = 0; i + 1 < c->evs.size(); i += 2) {
    HIPCHK(hipEventSynchronize(c->evs[i + 1]));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->evs[i], c->evs[i + 1]));
    c->kernel_ms += ms;
    c->kernel_launches++;
    (void)hipEventDestroy(c->evs[i]);
    (void)hipEventDestroy(c->evs[i + 1]);
  }
  c->evs.clear();
  return 0;
}

static int timing_mark(dasp_cent *c, hipStream_t st) {
  if (!c->timing) return 0;
  hipEvent_t e;
  HIPCHK(hipEventCreate(&e));
  HIPCHK(hipEventRecord(e, st));
  c->evs.push_back(e);
  if (c->evs.size() >= 512 && c->evs.size() % 2 == 0) return harvest_events(c);
  return 0;
}

// ------------------------------------------------------------------------------------------------
// derived tables
// ------------------------------------------------------------------------------------------------
static int set_spot(dasp_cent *c, const dasp_cent_array &sp) {
  // setSpotPsf (cmod/src/centmodule.c:2255-2301)
  c->spotDim = 0;
  c->psf = c->clip;
  c->pbf = 1;
  c->padded = c->clip;
  c->spot.clear();
  if (!sp.data) return 0;
  if (sp.ndim == 2) {
    c->spotDim = 2;
    c->psf = (int)sp.dims[0];
    if (sp.dims[1] != sp.dims[0]) return fail(DASP_CENT_EINVAL, "spotpsf must be square");
  } else if (sp.ndim == 4) {
    c->spotDim = 4;
    c->psf = (int)sp.dims[2];
    if (sp.dims[0] * sp.dims[1] != c->nsubaps || sp.dims[3] != sp.dims[2])
      return fail(DASP_CENT_EINVAL, "spotpsf should be (nsubx,nsubx,P,P)");
  } else {
    return fail(DASP_CENT_EINVAL, "centmodule - spotpsf not 2d or 4d");
  }
  c->pbf = c->preBinOrig;
  c->padded = c->pbf <= 1 ? c->psf : c->F;
  if (c->pbf < 1) c->pbf = 1;
  if (c->padded % c->pbf != 0)
    return fail(DASP_CENT_EINVAL, "paddedsize(%d) modulo preBinningFactor(%d) should be zero", c->padded, c->pbf);
  if (c->pbf > 1 && c->psf > c->padded)
    return fail(DASP_CENT_EINVAL, "spotpsf dimensions should be equal to or less than paddedsize (%d)", c->padded);
  if (c->psf < c->clip) return fail(DASP_CENT_EINVAL, "spotpsf dimensions should be at least clipsize");
  if (!is_pow2(c->psf) || c->psf > kMaxSide)
    return fail(DASP_CENT_EUNSUPPORTED, "spot kernel size %d: the device convolution needs a power of two <= %d", c->psf, kMaxSide);
  if (c->padded > kMaxSide) return fail(DASP_CENT_EUNSUPPORTED, "paddedsize %d > %d", c->padded, kMaxSide);
  const float *src = (const float *)sp.data;
  c->spot.assign(src, src + numel(sp));
  // OTF = r2c(flipArray(spot, mirror=1)) (prepareSpotPsf :2348-2372), stored as the full
  // Hermitian P x P spectrum for the device's complex convolution.
  const int P = c->psf;
  const int nk = c->spotDim == 2 ? 1 : c->nsubaps;
  std::vector<float2> otf((size_t)nk * P * P);
  std::vector<float> tmp(P * P);
  std::vector<cd> z(P * P);
  for (int k = 0; k < nk; ++k) {
    std::memcpy(tmp.data(), src + (long)k * P * P, sizeof(float) * P * P);
    host_flip_array(tmp.data(), P, 1);
    for (int q = 0; q < P * P; ++q) z[q] = cd(tmp[q], 0);
    host_dft2(z.data(), P, P, -1);
    for (int q = 0; q < P * P; ++q) otf[(size_t)k * P * P + q] = make_float2((float)z[q].real(), (float)z[q].imag());
  }
  if (upload(c->d_otf, otf.data(), otf.size() * sizeof(float2)) != hipSuccess)
    return fail(DASP_CENT_EHIP, "upload of spot OTFs failed");
  if (P == c->F && P % 2 == 0 && c->F == 2 * c->nimg) {
    // Table of the fused spot convolution (spot_convolve in fused_item.h), K = conj(OTF), for a
    // detector binned 2 x 2: the block sums over pixel pairs (2m, 2m+1) of an unnormalised
    // inverse transform are the P/2-point inverse transform of Y[u] f(u) + Y[u + P/2] f(u + P/2),
    // f(u) = 1 + exp(2 pi i u / P), in both axes, so both fold factors join K:
    // rows m = 1..P/2-1: K(u, m) f(u) f(m); rows 0 and P/2 (the packed column pair 0, P/2, whose
    // second member vanishes since f(P/2) = 0): K(u, 0) f(u) f(0) / 2 = K(u, 0) f(u).
    const int Hh = P / 2;
    std::vector<float2> tab((size_t)nk * (Hh + 1) * P);
    auto fold = [&](int x) { return cd(1.0 + std::cos(2.0 * M_PI * x / P), std::sin(2.0 * M_PI * x / P)); };
    for (int k = 0; k < nk; ++k) {
      const float2 *o = otf.data() + (size_t)k * P * P;
      float2 *t = tab.data() + (size_t)k * (Hh + 1) * P;
      for (int u = 0; u < P; ++u) {
        const cd k0 = cd(o[u * P].x, -o[u * P].y) * fold(u);
        t[u] = make_float2((float)k0.real(), (float)k0.imag());
        t[(size_t)Hh * P + u] = t[u];
        for (int m = 1; m < Hh; ++m) {
          const cd km = cd(o[u * P + m].x, -o[u * P + m].y) * fold(u) * fold(m);
          t[(size_t)m * P + u] = make_float2((float)km.real(), (float)km.imag());
        }
      }
    }
    if (upload(c->d_otfk, tab.data(), tab.size() * sizeof(float2)) != hipSuccess)
      return fail(DASP_CENT_EHIP, "upload of spot kernel table failed");
  } else {
    c->d_otfk.release();
  }
  return 0;
}

// updateTotSigArr (cmod/src/centmodule.c:2403-2427), float arithmetic as the reference.
static void compute_totsig(dasp_cent *c) {
  const int n2 = c->n * c->n;
  c->totsig.assign(c->nsubaps, 0.f);
  for (int i = 0; i < c->nsubaps; ++i) {
    float t = 0.f;
    for (int j = 0; j < n2; ++j) t += c->pupfn[(long)i * n2 + j];
    t *= (float)(c->F * c->F);
    if (c->spotDim) {
      const int P2 = c->psf * c->psf;
      const float *sp = c->spot.data() + (c->spotDim == 4 ? (long)i * P2 : 0);
      float tot = 0.f;
      for (int j = 0; j < P2; ++j) tot += sp[j];
      t *= tot * P2;
    }
    c->totsig[i] = t;
  }
}

static int upload_tables(dasp_cent *c) {
  if (!c->tablesDirty) return 0;
  const int S = c->nsubaps;
  std::vector<float> scale(S), sigv(S), skyadd(S);
  std::vector<unsigned char> pois(S);
  for (int i = 0; i < S; ++i) {
    const float nph = c->frac[i];
    sigv[i] = c->haveSigArr ? c->sigArr[i] : c->sig * nph;
    scale[i] = c->totsig[i] > 0.f ? (c->haveSigArr ? c->sigArr[i] / c->totsig[i] : c->sig * nph / c->totsig[i]) : 1.f;
    const float sky = c->skybrightness * nph;
    skyadd[i] = sky > 0.f ? sky : 0.f;
    pois[i] = (c->addPoisson == 1 && c->calsource == 0) && (c->totsig[i] > 0.f || sky > 0.f || c->haveSkyArr);
  }
  c->allScalePos = true;
  for (int i = 0; i < S; ++i) c->allScalePos = c->allScalePos && scale[i] > 0.f;
  if (upload(c->d_scale, scale.data(), S * 4) || upload(c->d_sigv, sigv.data(), S * 4) ||
      upload(c->d_skyadd, skyadd.data(), S * 4) || upload(c->d_poisson, pois.data(), S) ||
      upload(c->d_totsig, c->totsig.data(), S * 4))
    return fail(DASP_CENT_EHIP, "upload of per-subaperture tables failed");
  // binning operator A (nimg x clip) from prepareBinImage/binImage (:565-651), float weights
  const int Q = c->clip, M = c->nimg;
  std::vector<float> A((size_t)M * Q, 0.f);
  if (Q == M) {
    for (int i = 0; i < M; ++i) A[i * Q + i] = 1.f;
  } else {
    std::vector<float> t(Q), nx(Q);
    std::vector<int> idx(Q), pos(Q);
    int cnt = 0;
    for (int i = 0; i < Q; ++i) {
      cnt += M;
      if (cnt >= Q) {
        const int ex = cnt - Q;
        t[i] = (M - ex) / (float)M;
        nx[i] = ex / (float)M;
        idx[i] = 1;
        cnt = ex;
      } else {
        t[i] = 1.f;
        nx[i] = 0.f;
        idx[i] = 0;
      }
    }
    int ii = 0;
    for (int r = 0; r < Q; ++r) {
      pos[r] = ii;
      ii += idx[r];
    }
    for (int r = 0; r < Q; ++r) {
      if (pos[r] < M) A[pos[r] * Q + r] += t[r];
      if (r >= 1 && pos[r - 1] + 1 < M) A[(pos[r - 1] + 1) * Q + r] += nx[r - 1];
    }
  }
  if (upload(c->d_binA, A.data(), A.size() * 4)) return fail(DASP_CENT_EHIP, "upload of binning table failed");
  // an integer factor gives the plain block sum (the general kernel then adds the blocks directly)
  c->binB = 0;
  if (Q % M == 0) {
    const int B = Q / M;
    bool block = true;
    for (int I = 0; I < M && block; ++I)
      for (int i = 0; i < Q && block; ++i) block = A[(size_t)I * Q + i] == (i / B == I ? 1.f : 0.f);
    if (block) c->binB = B;
  }
  // fit matrix for the effective fit size
  int fs = c->parabolicFit + c->gaussianFit;
  if (fs >= 0 && fs < 3) fs = 3;
  if (fs > 0 && fs != c->fitMxSize) {
    if (!host_fit_matrix(fs, c->fitMx)) return fail(DASP_CENT_EINVAL, "fit matrix for fitsize %d is singular", fs);
    c->fitMxSize = fs;
  }
  // computeFit's window (fitsize around the peak, clamped to the central ncen x ncen of the
  // corrsize x corrsize image it fits) has to lie inside that image: the reference reads past the
  // image otherwise (cmod/src/centmodule.c:1069-1086), here refused
  if (c->parabolicFit != 0 || c->gaussianFit > 0) {
    const int win = fs < 0 ? 3 : fs;
    if (win > c->ncen || c->ncen > c->corrsize)
      return fail(DASP_CENT_EINVAL, "fit window %d must fit inside ncen %d, and ncen inside the %d x %d image", win,
                  c->ncen, c->corrsize, c->corrsize);
  }
  c->tablesDirty = false;
  return 0;
}

static int set_corr_pattern(dasp_cent *c, const dasp_cent_array &a) {
  c->haveCorr = false;
  if (!a.data) return 0;
  if (a.ndim != 4 || a.dims[0] * a.dims[1] != c->nsubaps || a.dims[2] != a.dims[3])
    return fail(DASP_CENT_EINVAL, "corrPattern should be (nsubx,nsubx,C,C)");
  const int C = (int)a.dims[2];
  if (C > kMaxSide) return fail(DASP_CENT_EUNSUPPORTED, "corrPattern size %d > %d", C, kMaxSide);
  const float *src = (const float *)a.data;
  const size_t per = (size_t)C * C;
  if (upload(c->d_corrPat, src, per * c->nsubaps * 4)) return fail(DASP_CENT_EHIP, "corrPattern upload failed");
  if (is_pow2(C)) {
    std::vector<float2> K(per * c->nsubaps);
    std::vector<cd> k(per);
    for (int s = 0; s < c->nsubaps; ++s) {
      host_hc_to_spectrum(src + per * s, C, C, k.data());
      for (size_t q = 0; q < per; ++q) K[per * s + q] = make_float2((float)k[q].real(), (float)k[q].imag());
    }
    if (upload(c->d_corrK, K.data(), K.size() * sizeof(float2))) return fail(DASP_CENT_EHIP, "corr spectrum upload failed");
    // the two-image correlation kernel's table (C = 8, 16): per row u and column v < C/2 the pair
    // (alpha, beta) of W = alpha P[u] + beta conj(P[-u]): alpha = K[u][v], beta = 0 for v >= 1; for
    // v = 0 the packed column carries columns 0 and C/2 of the image as P = X0 + i XH, and
    // X0 K[u][0] + i XH K[u][C/2] = P (K0 + KH) / 2 + conj(P[-u]) (K0 - KH) / 2. One 16-byte load
    // per value for every lane.
    c->d_corrKp.release();
    if (C == 8 || C == 16) {
      const int H = C / 2;
      std::vector<float4> Kp((size_t)C * H * c->nsubaps);
      for (int s = 0; s < c->nsubaps; ++s) {
        host_hc_to_spectrum(src + per * s, C, C, k.data());
        for (int u = 0; u < C; ++u) {
          float4 *row = &Kp[((size_t)s * C + u) * H];
          const cd k0 = k[(size_t)u * C], kh = k[(size_t)u * C + H];
          const cd al = (k0 + kh) * 0.5, be = (k0 - kh) * 0.5;
          row[0] = make_float4((float)al.real(), (float)al.imag(), (float)be.real(), (float)be.imag());
          for (int v = 1; v < H; ++v) {
            const float2 kv = K[per * s + (size_t)u * C + v];
            row[v] = make_float4(kv.x, kv.y, 0.f, 0.f);
          }
        }
      }
      if (upload(c->d_corrKp, Kp.data(), Kp.size() * sizeof(float4))) return fail(DASP_CENT_EHIP, "corr spectrum upload failed");
    }
  } else {
    c->d_corrK.release();
    c->d_corrKp.release();
  }
  c->haveCorr = true;
  return 0;
}

static int set_cent_weight(dasp_cent *c, const dasp_cent_array &a) {
  // setCentWeight (cmod/src/centmodule.c:2227-2253)
  c->centWeightDim = 0;
  c->haveCentWeight = false;
  if (!a.data) return 0;
  if (a.ndim != 2 && a.ndim != 4) return fail(DASP_CENT_EINVAL, "centmodule - centWeight not 2d or 4d");
  c->centWeightDim = a.ndim;
  c->centWeightSize = (int)a.dims[a.ndim - 2];
  if (a.dims[a.ndim - 2] != c->corrsize || a.dims[a.ndim - 1] != c->corrsize)
    return fail(DASP_CENT_EINVAL, "centweight array not correct dimensions (need to be nsubx,nsubx,corrsize,corrsize)");
  if (upload(c->d_centWeight, a.data, numel(a) * 4)) return fail(DASP_CENT_EHIP, "centWeight upload failed");
  c->haveCentWeight = true;
  return 0;
}

static bool check_f32(const dasp_cent_array &a) { return a.data != nullptr; }

// Threshold code the kernels apply: 0 and 1 as readCCD (cmod/src/centmodule.c:785-817); 2 the
// relative floor (the engine option of dasp_cent_set_relative_floor, on top of threshType 0);
// -1 for every other threshType, which readCCD treats as unknown (:797, :813): no read noise,
// no readBg and no floor.
static int kthresh(const dasp_cent *c) {
  if (c->threshType == 0) return c->relFloor ? 2 : 0;
  return c->threshType == 1 ? 1 : -1;
}


// ------------------------------------------------------------------------------------------------
// kernel argument assembly and launch
// ------------------------------------------------------------------------------------------------
static void fill_back(dasp_cent *c, BackArgs &b, int nframes) {
  std::memset(&b, 0, sizeof(b));
  const int M2 = c->nimg * c->nimg;
  b.active = c->d_active.as<int>();
  b.nactive = (int)c->active.size();
  b.nitems = (long)nframes * b.nactive;
  b.nsubx = c->nsubx;
  b.nsubaps = c->nsubaps;
  b.n = c->n;
  b.F = c->F;
  b.clip = c->clip;
  b.nimg = c->nimg;
  b.ncen = c->ncen;
  b.psf = c->psf;
  b.padded = c->padded;
  b.pbf = c->spotDim ? c->pbf : 1;
  b.corrsize = c->corrsize;
  b.npup = c->npup;
  b.otf = c->spotDim ? c->d_otf.as<float2>() : nullptr;
  b.otf_per_subap = c->spotDim == 4;
  b.binA = c->d_binA.as<float>();
  b.binB = c->binB;
  b.scale = c->d_scale.as<float>();
  b.totsig = c->d_totsig.as<float>();
  b.sigv = c->d_sigv.as<float>();
  b.pup = c->d_pup.as<float>();
  b.skyadd = c->d_skyadd.as<float>();
  b.skyarr = c->haveSkyArr ? c->d_skyarr.as<float>() : nullptr;
  b.poisson_on = c->d_poisson.as<unsigned char>();
  b.noise_stride = (long)c->nsubaps * M2;
  b.readnoise = c->calsource == 0 ? c->readnoise : 0.f;
  b.readbg = c->readbg;
  b.floor = c->noiseFloor;
  b.threshType = kthresh(c);
  b.opticalBinning = c->opticalBinning;
  b.calsource = c->calsource;
  b.useBrightest = c->useBrightest;
  b.useBrightestArr = c->haveUBArr ? c->d_ubArr.as<int>() : nullptr;
  b.pxlPower = c->pxlPower;
  b.imageOnly = c->imageOnly;
  b.corrMode = c->haveCorr ? c->correlationCentroiding : 0;
  if (c->correlationCentroiding == 5 && !c->haveCorr) b.corrMode = 0;
  b.corrThresh = c->corrThresh;
  b.corrK = (c->haveCorr && c->correlationCentroiding == 1) ? c->d_corrK.as<float2>() : nullptr;
  b.corrPattern = c->haveCorr ? c->d_corrPat.as<float>() : nullptr;
  b.centWeight = c->haveCentWeight ? c->d_centWeight.as<float>() : nullptr;
  b.centWeightDim = c->haveCentWeight ? c->centWeightDim : 0;
  {
    int fs = c->parabolicFit + c->gaussianFit;
    b.fitsize = fs < 0 ? -1 : (fs < 3 ? 3 : fs);
  }
  b.doFit = (c->parabolicFit != 0 || c->gaussianFit > 0);
  b.gaussianFit = c->gaussianFit;
  b.gaussianMinVal = c->gaussianMinVal;
  b.gaussianReplaceVal = c->gaussianReplaceVal;
  std::memcpy(b.fitMx, c->fitMx, sizeof(b.fitMx));
  b.calData = c->haveCalData ? c->d_calData.as<float>() : nullptr;
  b.calBounds = c->d_calBounds.as<int>();
  b.calSteps = c->d_calSteps.as<float>();
  b.calNSteps = c->calNSteps;
  b.calCoeff = c->haveCalCoeff ? c->d_calCoeff.as<float>() : nullptr;
  b.calNCoeff = c->calNCoeff;
  b.refCents = c->haveRef ? c->d_refCents.as<float>() : nullptr;
  b.cents_stride = (long)c->nsubaps * 2;
  b.bimg_stride = (long)c->nsubaps * M2;
  b.corr_stride = (long)c->nsubaps * c->corrsize * c->corrsize;
}

static bool need_pois(dasp_cent *c) { return c->addPoisson == 1 && c->calsource == 0; }
static bool need_read(dasp_cent *c) { return c->readnoise != 0.f && c->calsource == 0; }

// The fused kernel covers: no spot kernel, clip == fftsize, integer binning of a shape it was
// instantiated for, phase only, plain or weighted CoG without calibration; the brightest-pixel
// selection in the FP32 kernels only.
static bool fused_spot(dasp_cent *c) {
  return c->spotDim != 0 && c->psf == c->F && c->padded == c->F && c->pbf == 1 && c->F == 2 * pad_pow2(c->n) &&
         c->d_otfk.p != nullptr;
}

// Centroid stage run by cent_corr.hip after the fused kernels: FFT correlation (mode 1) with a
// fit or a centre-on-brightest CoG, or a fit on the image itself.
static bool corr_stage(dasp_cent *c) {
  const bool fit = c->parabolicFit != 0 || c->gaussianFit > 0;
  return (c->correlationCentroiding == 1 && c->haveCorr) || (c->correlationCentroiding == 0 && fit);
}

static bool corr_stage_ok(dasp_cent *c) {
  if (c->correlationCentroiding == 0) return !corr_stage(c) || corr_fit_supported(c->nimg);
  return c->correlationCentroiding == 1 && c->haveCorr && corr_fit_supported(c->corrsize) && c->corrsize >= c->nimg;
}

static bool fused_options(dasp_cent *c) {
  return (c->spotDim == 0 || fused_spot(c)) && c->clip == c->F && c->atmosPhaseType == 0 && c->opticalBinning == 0 &&
         c->imageOnly == 0 && corr_stage_ok(c) &&
         !c->haveCalData && !c->haveCalCoeff;
}

// selectBrightestPixels runs only without calsource (cmod/src/centmodule.c:1883-1884).
static bool brightest_on(dasp_cent *c) { return c->calsource == 0 && (c->useBrightest != 0 || c->haveUBArr); }

// The DFT-GEMM kernels cover the fused options (no brightest-pixel selection) for n = 16 / fftsize 32 and n in 17..32 /
// fftsize 64, nimg 16.
static bool is_dft_form(int form) {
  return form == DASP_CENT_FORM_DFT_FP16 || form == DASP_CENT_FORM_DFT_BF16 || form == DASP_CENT_FORM_DFT_FP16_1;
}

static bool use_dft(dasp_cent *c) {
  return is_dft_form(c->form) && fused_options(c) && !brightest_on(c) && c->spotDim == 0 &&
         dft_supported(c->n, c->F, c->nimg);
}

static bool use_fused(dasp_cent *c) {
  return c->form != DASP_CENT_FORM_FFT32_GENERAL && fused_options(c) &&
         (use_dft(c) || fused_supported(pad_pow2(c->n), c->F, c->nimg, c->spotDim != 0));
}

// Arguments of the fused kernels (FP32 and DFT-GEMM forms) for nframes device frames.
static void fused_args(dasp_cent *c, const float *d_phs, long phs_stride, int nframes, float *d_cents,
                       float *d_bimg, const float *d_zpois, const float *d_zread, FusedArgs &fa) {
  const long nitems = (long)nframes * (long)c->active.size();
  fa = FusedArgs{};
  fa.phs = d_phs;
  fa.phs_stride = phs_stride;
  fa.pup = c->d_pup.as<float>();
  fa.tilt = c->d_tilt.as<float>();
  fa.tiltstep = (float)(M_PI / c->F * (c->F + 1.));
  fa.fast_trig = c->fastTrig;
  fa.active = c->d_active.as<int>();
  fa.winoff = c->d_winoff.as<int>();
  fa.nactive = (int)c->active.size();
  fa.nitems = nitems;
  fa.nsubx = c->nsubx;
  fa.n = c->n;
  fa.npup = c->npup;
  fa.nimg = c->nimg;
  fa.ncen = c->ncen;
  fa.scale = c->d_scale.as<float>();
  fa.skyadd = c->d_skyadd.as<float>();
  fa.skyarr = c->haveSkyArr ? c->d_skyarr.as<float>() : nullptr;
  fa.poisson_on = c->d_poisson.as<unsigned char>();
  fa.zpois = d_zpois;
  fa.zread = d_zread;
  fa.noise_stride = (long)c->nsubaps * c->nimg * c->nimg;
  fa.readnoise = c->calsource == 0 ? c->readnoise : 0.f;
  fa.readbg = c->readbg;
  fa.floor = c->noiseFloor;
  fa.threshType = kthresh(c);
  fa.calsource = c->calsource;
  fa.pxlPower = c->pxlPower;
  fa.centWeight = c->haveCentWeight ? c->d_centWeight.as<float>() : nullptr;
  fa.centWeightDim = c->haveCentWeight ? c->centWeightDim : 0;
  fa.refCents = c->haveRef ? c->d_refCents.as<float>() : nullptr;
  fa.cents = d_cents;
  fa.cents_stride = (long)c->nsubaps * 2;
  fa.bimg = d_bimg;
  fa.bimg_stride = (long)c->nsubaps * c->nimg * c->nimg;
  fa.slow_detector = (d_zpois && need_pois(c)) || (d_zread && need_read(c)) || c->haveSkyArr ||
                     kthresh(c) == 2 || (c->pxlPower != 1.f && c->pxlPower != 0.f) || c->haveCentWeight ||
                     brightest_on(c);
  fa.useBrightest = brightest_on(c) && !c->haveUBArr ? c->useBrightest : 0;
  fa.useBrightestArr = brightest_on(c) && c->haveUBArr ? c->d_ubArr.as<int>() : nullptr;
  {
    // the fast detector's pixel is scale * bin + (sky + readbg), then the threshold of type tt
    const int tt = fa.threshType;
    const bool adds = c->skybrightness > 0.f || (tt >= 0 && tt <= 2 && c->readbg != 0.f);
    const bool thr_noop = !(tt == 0 || tt == 1) || (tt == 0 && c->noiseFloor == 0.f) || (tt == 1 && c->noiseFloor <= 0.f);
    fa.bins_only = !fa.slow_detector && !d_bimg && !adds && thr_noop && c->ncen == c->nimg && c->allScalePos;
  }
  fa.otfk = c->spotDim ? c->d_otfk.as<float2>() : nullptr;
  fa.otf_per_subap = c->spotDim == 4;
}

// Device scratch of one phase -> slopes call: images and correlation images the kernels need but
// the caller did not provide, the resampled phase grid and the general path's hll.
struct Scratch {
  float *bimg = nullptr, *corr = nullptr, *interp = nullptr, *hll = nullptr;
};

// Phase -> centroids for nframes frames already resident on the device. d_bimg / d_corrimg are
// the images the caller gets back (or the scratch standing in for them).
static int launch_phase(dasp_cent *c, const float *d_phs, long phs_stride, int nframes, float *d_cents,
                        float *d_bimg, float *d_corrimg, const float *d_zpois, const float *d_zread,
                        const Scratch &sc, hipStream_t st) {
  int rc = upload_tables(c);
  if (rc) return rc;
  if (c->interpolate) {   // interpolatePhase (:1628-1758) onto the nsubx*phasesize grid
    const long tot2 = (long)c->npup * c->npup;
    if (!sc.interp) return fail(DASP_CENT_EINVAL, "internal: no interpolation scratch");
    InterpArgs ia{};
    ia.in = d_phs;
    ia.in_stride = phs_stride;
    ia.m = c->inputPhaseSize;
    ia.out = sc.interp;
    ia.tot = c->npup;
    ia.scale = (float)((c->inputPhaseSize - 1.) / (c->npup - 1));
    ia.nframes = nframes;
    HIPCHK(launch_interp(ia, st));
    d_phs = sc.interp;
    phs_stride = tot2;
  }
  const long nitems = (long)nframes * (long)c->active.size();
  const int np = pad_pow2(c->n);
  const float *amp = c->atmosPhaseType == 1 ? d_phs + (long)c->npup * c->npup : nullptr;
  bool inactive_done = false;
  if (nitems > 0 && use_fused(c)) {
    FusedArgs fa;
    fused_args(c, d_phs, phs_stride, nframes, d_cents, d_bimg, d_zpois, d_zread, fa);
    const bool stage = corr_stage(c);
    if (stage && !d_bimg) return fail(DASP_CENT_EINVAL, "the correlation/fit stage needs the image buffer");
    fa.no_cents = stage;
    if (!c->inactive.empty() && c->imageOnly == 0 && d_cents) {   // zeroed by the image kernel
      fa.inactive = c->d_inactive.as<int>();
      fa.ninactive = (int)c->inactive.size();
      inactive_done = true;
    }
    CorrArgs ca{};
    const bool corr = c->correlationCentroiding == 1;
    const int N = corr ? c->corrsize : c->nimg;
    if (stage) {
      ca.img = d_bimg;
      ca.img_stride = (long)c->nsubaps * c->nimg * c->nimg;
      ca.active = c->d_active.as<int>();
      ca.nactive = (int)c->active.size();
      ca.nitems = nitems;
      ca.nimg = c->nimg;
      ca.ncen = c->ncen;
      ca.corrK = corr ? c->d_corrK.as<float2>() : nullptr;
      ca.corrKp = corr ? c->d_corrKp.as<float4>() : nullptr;
      ca.corrThresh = c->corrThresh;
      ca.doFit = c->parabolicFit != 0 || c->gaussianFit > 0;
      const int fs = c->parabolicFit + c->gaussianFit;
      ca.fitsize = fs < 0 ? -1 : (fs < 3 ? 3 : fs);
      ca.gaussianFit = c->gaussianFit;
      ca.gaussianMinVal = c->gaussianMinVal;
      ca.gaussianReplaceVal = c->gaussianReplaceVal;
      std::memcpy(ca.fitMx, c->fitMx, sizeof(ca.fitMx));
      ca.centWeight = c->haveCentWeight ? c->d_centWeight.as<float>() : nullptr;
      ca.centWeightDim = c->haveCentWeight ? c->centWeightDim : 0;
      ca.refCents = c->haveRef ? c->d_refCents.as<float>() : nullptr;
      ca.cents = d_cents;
      ca.cents_stride = (long)c->nsubaps * 2;
      ca.corrimg = corr ? d_corrimg : nullptr;
      ca.corr_stride = (long)c->nsubaps * N * N;
    }
    // With the correlation/fit stage and several frames, the frames go in chunks: the image
    // kernel of chunk k+1 (matrix-core and memory heavy) runs on st while the stage of chunk k
    // (vector heavy) runs on a side stream; the side stream joins st at the end. Not in the
    // allocation-free entry point (no stream or event creation there), and not while st is being
    // captured unless the side stream already exists.
    int nch = 1;
    if (stage && nframes >= 2 && !t_forbid_alloc) {
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      const bool capturing = hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
      if (!c->corrSide && !capturing) {
        if (hipStreamCreateWithFlags(&c->corrSide, hipStreamNonBlocking) != hipSuccess) c->corrSide = nullptr;
        if (c->corrSide && (hipEventCreateWithFlags(&c->corrFork, hipEventDisableTiming) != hipSuccess ||
                            hipEventCreateWithFlags(&c->corrJoin, hipEventDisableTiming) != hipSuccess)) {
          return fail(DASP_CENT_EHIP, "correlation side stream set-up failed");
        }
      }
      if (c->corrSide && c->corrFork && c->corrJoin) nch = nframes >= 16 ? 4 : (nframes >= 4 ? 2 : 1);
    }
    const long nact = (long)c->active.size();
    int trc = timing_mark(c, st);
    if (trc) return trc;
    for (int k = 0; k < nch; ++k) {
      const int f0 = (int)((long)nframes * k / nch), f1 = (int)((long)nframes * (k + 1) / nch);
      FusedArgs fk = fa;
      fk.phs = fa.phs + (long)f0 * fa.phs_stride;
      fk.nitems = (long)(f1 - f0) * nact;
      if (fk.cents) fk.cents = fa.cents + (long)f0 * fa.cents_stride;
      if (fk.bimg) fk.bimg = fa.bimg + (long)f0 * fa.bimg_stride;
      if (fk.zpois) fk.zpois = fa.zpois + (long)f0 * fa.noise_stride;
      if (fk.zread) fk.zread = fa.zread + (long)f0 * fa.noise_stride;
      if (use_dft(c))
        HIPCHK(launch_dft(fk, c->form == DASP_CENT_FORM_DFT_BF16, c->form != DASP_CENT_FORM_DFT_FP16_1, st, c->gridImage));
      else HIPCHK(launch_fused(fk, np, c->F, c->spotDim != 0, st, c->gridImage));
      if (nch > 1) {
        CorrArgs ck = ca;
        ck.img = ca.img + (long)f0 * ca.img_stride;
        ck.nitems = (long)(f1 - f0) * nact;
        ck.cents = ca.cents + (long)f0 * ca.cents_stride;
        if (ck.corrimg) ck.corrimg = ca.corrimg + (long)f0 * ca.corr_stride;
        HIPCHK(hipEventRecord(c->corrFork, st));
        HIPCHK(hipStreamWaitEvent(c->corrSide, c->corrFork, 0));
        HIPCHK(launch_corr_fit(ck, N, corr, c->corrSide, c->gridCorr));
      }
    }
    trc = timing_mark(c, st);   // the timed interval is the image kernel alone
    if (trc) return trc;
    if (nch > 1) {
      HIPCHK(hipEventRecord(c->corrJoin, c->corrSide));
      HIPCHK(hipStreamWaitEvent(st, c->corrJoin, 0));
    } else if (stage) {
      HIPCHK(launch_corr_fit(ca, N, corr, st, c->gridCorr));
    }
  } else if (nitems > 0) {
    const long hll_elems = nitems * (long)c->padded * c->padded;
    if (!sc.hll) return fail(DASP_CENT_EINVAL, "internal: no image-formation scratch");
    if (c->padded > c->F) HIPCHK(launch_fill(sc.hll, hll_elems, 0.f, st));
    FormArgs fa{};
    fa.phs = d_phs;
    fa.amp = amp;
    fa.phs_stride = phs_stride;
    fa.pup = c->d_pup.as<float>();
    fa.tilt = c->d_tilt.as<float>();
    fa.active = c->d_active.as<int>();
    fa.nactive = (int)c->active.size();
    fa.nitems = nitems;
    fa.nsubx = c->nsubx;
    fa.n = c->n;
    fa.npup = c->npup;
    fa.padded = c->padded;
    fa.hll = sc.hll;
    fa.fast_trig = c->fastTrig ? 1 : 0;
    int trc = timing_mark(c, st);
    if (trc) return trc;
    HIPCHK(launch_form(fa, np, c->F, st));
    trc = timing_mark(c, st);
    if (trc) return trc;
    BackArgs b;
    fill_back(c, b, nframes);
    b.hll = sc.hll;
    b.amp = amp;
    b.phs_stride = phs_stride;
    b.zpois = d_zpois;
    b.zread = d_zread;
    b.cents = d_cents;
    b.bimg = d_bimg;
    b.corrimg = d_corrimg;
    HIPCHK(launch_backend(b, st));
  }
  if (!inactive_done && !c->inactive.empty() && (c->imageOnly == 0 ? d_cents != nullptr : d_bimg != nullptr)) {
    InactiveArgs ia{};
    ia.inactive = c->d_inactive.as<int>();
    ia.ninactive = (int)c->inactive.size();
    ia.nframes = nframes;
    ia.nsubaps = c->nsubaps;
    ia.nimg = c->nimg;
    ia.imageOnly = c->imageOnly;
    ia.readnoise = c->calsource == 0 ? c->readnoise : 0.f;
    ia.readbg = c->readbg;
    ia.floor = c->noiseFloor;
    ia.threshType = kthresh(c);
    ia.zread = d_zread;
    ia.noise_stride = (long)c->nsubaps * c->nimg * c->nimg;
    ia.cents = d_cents;
    ia.cents_stride = (long)c->nsubaps * 2;
    ia.bimg = d_bimg;
    ia.bimg_stride = (long)c->nsubaps * c->nimg * c->nimg;
    HIPCHK(launch_inactive(ia, st));
  }
  return 0;
}

// Draw the readout's noise streams (seed, 2k) and (seed, 2k+1) on the device, in stream order
// before the image kernels (the host generator's sequence, without a host round trip).
static int draw_noise(dasp_cent *c, bool pois, bool rd) {
  const long per = (long)c->nsubaps * c->nimg * c->nimg;
  const unsigned long k = c->nreadouts;
  if (pois) {
    HIPCHK(c->d_zpois.ensure(per * 4));
    HIPCHK(launch_normals(c->seed, 2 * k, per, c->d_zpois.as<float>(), nullptr));
  }
  if (rd) {
    HIPCHK(c->d_zread.ensure(per * 4));
    HIPCHK(launch_normals(c->seed, 2 * k + 1, per, c->d_zread.as<float>(), nullptr));
  }
  return 0;
}

static double now_s() {
  struct timeval t;
  gettimeofday(&t, nullptr);
  return t.tv_sec + 1e-6 * t.tv_usec;
}

// Copy device outputs into the borrowed host arrays with the reference's write pattern.
static int copy_back(dasp_cent *c) {
  const int M2 = c->nimg * c->nimg;
  if (c->interpolate && c->h_outphase) {   // interpolatePhase writes each active window (:1748-1755)
    const size_t tot2 = (size_t)c->npup * c->npup;
    c->h_stage.resize(tot2);
    HIPCHK(hipMemcpy(c->h_stage.data(), c->d_interp.p, tot2 * 4, hipMemcpyDeviceToHost));
    for (int s : c->active) {
      const size_t o = (size_t)(s / c->nsubx) * c->n * c->npup + (size_t)(s % c->nsubx) * c->n;
      for (int y = 0; y < c->n; ++y)
        std::memcpy(c->h_outphase + o + (size_t)y * c->npup, c->h_stage.data() + o + (size_t)y * c->npup, c->n * 4);
    }
  }
  if (c->imageOnly == 0 && c->h_cents)
    HIPCHK(hipMemcpy(c->h_cents, c->d_cents.p, (size_t)c->nsubaps * 2 * 4, hipMemcpyDeviceToHost));
  if (c->h_bimg) {
    const size_t nb = (size_t)c->nsubaps * M2 * 4;
    if (c->imageOnly) {
      HIPCHK(hipMemcpy(c->h_bimg, c->d_bimg.p, nb, hipMemcpyDeviceToHost));
    } else {
      c->h_stage.resize((size_t)c->nsubaps * M2);
      HIPCHK(hipMemcpy(c->h_stage.data(), c->d_bimg.p, nb, hipMemcpyDeviceToHost));
      for (int s : c->active) std::memcpy(c->h_bimg + (size_t)s * M2, c->h_stage.data() + (size_t)s * M2, M2 * 4);
    }
  }
  if (c->h_corrimg && c->haveCorr && c->correlationCentroiding && c->imageOnly == 0) {
    const int C2 = c->corrsize * c->corrsize;
    c->h_stage.resize((size_t)c->nsubaps * C2);
    HIPCHK(hipMemcpy(c->h_stage.data(), c->d_corrimg.p, (size_t)c->nsubaps * C2 * 4, hipMemcpyDeviceToHost));
    for (int s : c->active) std::memcpy(c->h_corrimg + (size_t)s * C2, c->h_stage.data() + (size_t)s * C2, C2 * 4);
  }
  return 0;
}

// ------------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------------
// ------------------------------------------------------------------------------------------------
// device-resident runs
// ------------------------------------------------------------------------------------------------
static int check_device_run(dasp_cent *c, const float *d_phs, long phs_stride, int nframes, const float *d_cents,
                            const float *d_zpois, const float *d_zread) {
  if (!c) return fail(DASP_CENT_EINVAL, "null handle");
  if (!d_phs || nframes < 1 || !d_cents) return fail(DASP_CENT_EINVAL, "d_phs, d_cents and nframes >= 1 are required");
  const long side = c->interpolate ? c->inputPhaseSize : c->npup;
  if (nframes > 1 && phs_stride < side * side * (c->atmosPhaseType + 1)) return fail(DASP_CENT_EINVAL, "phs_stride too small");
  if (need_pois(c) && !d_zpois) return fail(DASP_CENT_EINVAL, "Poisson noise is on: d_zpois is required");
  if (need_read(c) && !d_zread) return fail(DASP_CENT_EINVAL, "read noise is on: d_zread is required");
  // the detectors read four consecutive normals per 16-byte load
  if (((uintptr_t)d_zpois | (uintptr_t)d_zread) % 16 != 0)
    return fail(DASP_CENT_EINVAL, "d_zpois and d_zread must be 16-byte aligned");
  return 0;
}

// Bytes of each scratch buffer a call over nframes needs (0 when not needed): images when the
// caller passes none but the general path or the correlation stage reads them, correlation images
// likewise, the resampled phase grid, and the general path's hll.
struct ScratchSizes {
  size_t bimg = 0, corr = 0, interp = 0, hll = 0;
};

static ScratchSizes scratch_sizes(dasp_cent *c, int nframes, bool have_bimg, bool have_corr) {
  ScratchSizes z;
  const size_t F = (size_t)nframes;
  if (!have_bimg && (!use_fused(c) || corr_stage(c))) z.bimg = F * c->nsubaps * c->nimg * c->nimg * 4;
  if (!have_corr && c->correlationCentroiding && c->haveCorr) z.corr = F * c->nsubaps * c->corrsize * c->corrsize * 4;
  if (c->interpolate) z.interp = F * c->npup * c->npup * 4;
  if (!use_fused(c) && !c->active.empty()) z.hll = F * c->active.size() * c->padded * c->padded * 4;
  return z;
}

// Scratch from the handle's own buffers (grown as needed).
static int device_scratch(dasp_cent *c, int nframes, float *bimg, float *corr, Scratch &sc) {
  const ScratchSizes z = scratch_sizes(c, nframes, bimg != nullptr, corr != nullptr);
  if (z.bimg) HIPCHK(grow_scratch(c, c->d_bscratch, z.bimg));
  if (z.corr) HIPCHK(grow_scratch(c, c->d_cscratch, z.corr));
  if (z.interp) HIPCHK(grow_scratch(c, c->d_interp, z.interp));
  if (z.hll) HIPCHK(grow_scratch(c, c->d_hll, z.hll));
  sc.bimg = bimg ? bimg : (z.bimg ? c->d_bscratch.as<float>() : nullptr);
  sc.corr = corr ? corr : (z.corr ? c->d_cscratch.as<float>() : nullptr);
  sc.interp = z.interp ? c->d_interp.as<float>() : nullptr;
  sc.hll = z.hll ? c->d_hll.as<float>() : nullptr;
  return 0;
}

// Scratch carved from the fixed device array (256-byte aligned pieces); fails when it is too small.
static int fixed_scratch(dasp_cent *c, int nframes, float *bimg, float *corr, Scratch &sc) {
  const ScratchSizes z = scratch_sizes(c, nframes, bimg != nullptr, corr != nullptr);
  auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
  const size_t total = up(z.bimg) + up(z.corr) + up(z.interp) + up(z.hll);
  if (total > kNoAllocScratchBytes)
    return fail(DASP_CENT_EINVAL,
                "run_device_noalloc: %d frames need %zu bytes of scratch, more than the fixed %zu "
                "(pass image/correlation buffers or fewer frames)",
                nframes, total, (size_t)kNoAllocScratchBytes);
  if (total > 0 && !c->noallocScratch) return fail(DASP_CENT_EHIP, "run_device_noalloc: no scratch array on this device");
  char *p = (char *)c->noallocScratch;
  size_t off = 0;
  auto take = [&](size_t nb) -> float * {
    if (!nb) return nullptr;
    float *q = (float *)(p + off);
    off += up(nb);
    return q;
  };
  sc.bimg = bimg ? bimg : take(z.bimg);
  sc.corr = corr ? corr : take(z.corr);
  sc.interp = take(z.interp);
  sc.hll = take(z.hll);
  return 0;
}

// ------------------------------------------------------------------------------------------------
// multi-WFS step
// ------------------------------------------------------------------------------------------------
namespace {

struct BatchLaunch {
  int np = 0, F = 0, nimg = 0, fast = 0;
  bool spot = false;
  std::vector<int> items;      // indices into the batch's item list
  long groups = 0;
  DevBuf desc, counters;
  int waves = 0;
  long wgs = 0;
  int chunk = -1;              // -1: not chosen yet
  float ms = -1.f;
};

}  // namespace

struct dasp_cent_batch {
  std::vector<dasp_cent_batch_item> items;
  std::vector<unsigned long> gens;   // handle generations the graph was built on
  std::vector<unsigned long> epochs; // handle scratch epochs the graph was built on
  std::vector<BatchLaunch> launches;
  std::vector<int> others;           // items running their own kernels inside the graph
  DevBuf inactive;                   // InactiveDesc[] of the items in persistent launches
  int ninactive_desc = 0;
  long inactive_total = 0;
  int cus = 0;
  hipStream_t cap = nullptr;
  // side streams and events of the captured step's fork and join (default priority)
  static constexpr int kSide = 3;
  hipStream_t side[kSide] = {nullptr, nullptr, nullptr};
  hipEvent_t fork = nullptr, join[kSide] = {nullptr, nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static bool batch_fusable(dasp_cent *c) {
  return use_fused(c) && !use_dft(c) && !corr_stage(c) && !c->interpolate &&
         batch_supported(pad_pow2(c->n), c->F, c->nimg, c->spotDim != 0);
}

static void batch_default_config(dasp_cent_batch *b, BatchLaunch &L) {
  if (L.chunk < 0) L.chunk = 0;
  if (L.waves == 0) L.waves = 4;
  const int per_cu = batch_resident_wgs(L.np, L.F, L.nimg, L.spot, L.waves, L.fast);
  const long full = (L.groups + L.waves - 1) / L.waves;
  long w = (long)(per_cu > 0 ? per_cu : 1) * b->cus;
  if (L.wgs <= 0) L.wgs = w < full ? w : full;
}

static void batch_drop_graph(dasp_cent_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

static hipError_t batch_launch(const BatchLaunch &L, int ndesc, int waves, int chunk, long wgs, hipStream_t st) {
  BatchArgs a;
  a.desc = L.desc.as<BatchDesc>();
  a.ndesc = ndesc;
  a.total_groups = L.groups;
  a.counters = L.counters.as<unsigned long long>();
  return launch_batch(a, L.np, L.F, L.nimg, L.spot, waves, chunk, L.fast, wgs, st);
}

// Descriptors, queue counters and inactive-subaperture lists from the handles' current state.
// Tuned configurations of launches whose key survives are kept.
static int batch_layout(dasp_cent_batch *b) {
  std::vector<BatchLaunch> old;
  old.swap(b->launches);
  b->others.clear();
  b->gens.clear();
  std::vector<InactiveDesc> ina;
  long ina_total = 0;
  for (size_t i = 0; i < b->items.size(); ++i) {
    const dasp_cent_batch_item &it = b->items[i];
    dasp_cent *c = it.handle;
    b->gens.push_back(c->gen);
    int rc = check_device_run(c, it.phs, it.phs_stride, it.nframes, it.cents, it.zpois, it.zread);
    if (rc) return fail(rc, "batch item %zu: %s", i, g_err.c_str());
    if (c->inputDims != 0) return fail(DASP_CENT_EINVAL, "batch item %zu: image-input handles run with run_slope", i);
    rc = upload_tables(c);
    if (rc) return rc;
    if (!batch_fusable(c)) {
      b->others.push_back((int)i);
      continue;
    }
    const int np = pad_pow2(c->n), fast = c->fastTrig;
    const bool spot = c->spotDim != 0;
    BatchLaunch *L = nullptr;
    for (BatchLaunch &x : b->launches)
      if (x.np == np && x.F == c->F && x.nimg == c->nimg && x.fast == fast && x.spot == spot) L = &x;
    if (!L) {
      b->launches.emplace_back();
      L = &b->launches.back();
      L->np = np;
      L->F = c->F;
      L->nimg = c->nimg;
      L->fast = fast;
      L->spot = spot;
      for (BatchLaunch &o : old)
        if (o.np == np && o.F == c->F && o.nimg == c->nimg && o.fast == fast && o.spot == spot) {
          L->waves = o.waves;
          L->wgs = o.wgs;
          L->chunk = o.chunk;
          L->ms = o.ms;
        }
    }
    L->items.push_back((int)i);
    if (!c->inactive.empty()) {
      InactiveDesc q;
      q.inactive = c->d_inactive.as<int>();
      q.ninactive = (int)c->inactive.size();
      q.cents = it.cents;
      q.cents_stride = (long)c->nsubaps * 2;
      q.start = ina_total;
      ina.push_back(q);
      ina_total += (long)it.nframes * q.ninactive;
    }
  }
  for (BatchLaunch &o : old) {
    o.desc.release();
    o.counters.release();
  }
  for (BatchLaunch &L : b->launches) {
    std::vector<BatchDesc> d;
    const int spw = 64 / L.np;
    long g0 = 0;
    for (int i : L.items) {
      const dasp_cent_batch_item &it = b->items[i];
      BatchDesc q;
      fused_args(it.handle, it.phs, it.phs_stride, it.nframes, it.cents, it.bimg, it.zpois, it.zread, q.args);
      q.group0 = g0;
      g0 += (q.args.nitems + spw - 1) / spw;
      d.push_back(q);
    }
    L.groups = g0;
    if (upload(L.desc, d.data(), d.size() * sizeof(BatchDesc))) return fail(DASP_CENT_EHIP, "descriptor upload failed");
    HIPCHK(L.counters.ensure(2 * sizeof(unsigned long long)));
    HIPCHK(hipMemset(L.counters.p, 0, 2 * sizeof(unsigned long long)));
    batch_default_config(b, L);
  }
  for (int i : b->others) {
    dasp_cent_batch_item &it = b->items[i];
    Scratch sc;
    int rc = device_scratch(it.handle, it.nframes, it.bimg, it.corrimg, sc);
    if (rc) return rc;
  }
  // scratch addresses are final here: the graph captured next uses exactly these
  b->epochs.clear();
  for (const dasp_cent_batch_item &it : b->items) b->epochs.push_back(it.handle->scratchEpoch);
  b->ninactive_desc = (int)ina.size();
  b->inactive_total = ina_total;
  if (!ina.empty() && upload(b->inactive, ina.data(), ina.size() * sizeof(InactiveDesc)))
    return fail(DASP_CENT_EHIP, "descriptor upload failed");
  return 0;
}

// Enqueue one step: the persistent launches, the other WFS's own kernels, the inactive slopes.
// The tasks touch disjoint outputs and scratch, so with side streams (the captured step) the
// largest persistent launch stays on st and the others are forked onto the side streams and
// joined back: the graph holds them as parallel branches, and the small launches run beside the
// large one instead of after it.
static int batch_enqueue(dasp_cent_batch *b, hipStream_t st, bool fork) {
  std::vector<std::function<int(hipStream_t)>> tasks;
  size_t big = 0;
  for (size_t k = 0; k < b->launches.size(); ++k)
    if (b->launches[k].groups > b->launches[big].groups) big = k;
  for (size_t k = 0; k < b->launches.size(); ++k) {
    const BatchLaunch *L = &b->launches[k];
    auto t = [L](hipStream_t s) -> int {
      HIPCHK(batch_launch(*L, (int)L->items.size(), L->waves, L->chunk, L->wgs, s));
      return 0;
    };
    if (k == big) tasks.insert(tasks.begin(), t);
    else tasks.push_back(t);
  }
  for (int i : b->others) {
    tasks.push_back([b, i](hipStream_t s) -> int {
      dasp_cent_batch_item &it = b->items[i];
      dasp_cent *c = it.handle;
      Scratch sc;
      int rc = device_scratch(c, it.nframes, it.bimg, it.corrimg, sc);   // sized at layout time: no allocation here
      if (rc) return rc;
      const bool timing = c->timing;
      c->timing = false;   // no event records inside a captured step
      rc = launch_phase(c, it.phs, it.phs_stride, it.nframes, it.cents, sc.bimg, sc.corr, it.zpois, it.zread, sc, s);
      c->timing = timing;
      return rc;
    });
  }
  tasks.push_back([b](hipStream_t s) -> int {
    HIPCHK(launch_batch_inactive(b->inactive.as<InactiveDesc>(), b->ninactive_desc, b->inactive_total, s));
    return 0;
  });
  if (!fork || tasks.size() == 1) {
    for (auto &t : tasks) {
      int rc = t(st);
      if (rc) return rc;
    }
    return 0;
  }
  const int nside = (int)std::min<size_t>(dasp_cent_batch::kSide, tasks.size() - 1);
  HIPCHK(hipEventRecord(b->fork, st));
  for (int k = 0; k < nside; ++k) HIPCHK(hipStreamWaitEvent(b->side[k], b->fork, 0));
  int rc = tasks[0](st);
  for (size_t k = 1; k < tasks.size() && !rc; ++k) rc = tasks[k](b->side[(k - 1) % nside]);
  // join every side stream back (also after a failure, so that the capture can end cleanly)
  for (int k = 0; k < nside; ++k) {
    hipError_t e = hipEventRecord(b->join[k], b->side[k]);
    if (e == hipSuccess) e = hipStreamWaitEvent(st, b->join[k], 0);
    if (e != hipSuccess && !rc) rc = fail(DASP_CENT_EHIP, "step fork/join: %s", hipGetErrorString(e));
  }
  return rc;
}

static int batch_capture(dasp_cent_batch *b) {
  batch_drop_graph(b);
  HIPCHK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
  int rc = batch_enqueue(b, b->cap, true);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->cap, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(DASP_CENT_EHIP, "step capture failed: %s", hipGetErrorString(e));
  b->graph = g;
  HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return 0;
}

static int batch_refresh(dasp_cent_batch *b) {
  const size_t n = b->items.size();
  bool stale = !b->exec || b->gens.size() != n || b->epochs.size() != n;   // short after a failed layout
  for (size_t i = 0; i < b->items.size() && !stale; ++i)
    stale = b->items[i].handle->gen != b->gens[i] || b->items[i].handle->scratchEpoch != b->epochs[i];
  if (!stale) return 0;
  int rc = batch_layout(b);
  if (rc) {
    batch_drop_graph(b);   // never replay a graph whose layout could not be refreshed
    return rc;
  }
  return batch_capture(b);
}

extern "C" {

const char *dasp_cent_last_error(void) { return g_err.c_str(); }
int dasp_cent_abi_version(void) { return DASP_CENT_ABI_VERSION; }

int dasp_cent_draw_normals(unsigned long seed, unsigned long stream, long n, float *out) {
  if (n < 0 || (n > 0 && !out)) return fail(DASP_CENT_EINVAL, "bad arguments");
  host_draw_normals(seed, stream, n, out);
  return 0;
}

int dasp_cent_draw_normals_device(unsigned long seed, unsigned long stream, long n, float *d_out, void *hstream) {
  if (n < 0 || (n > 0 && !d_out)) return fail(DASP_CENT_EINVAL, "draw_normals_device: bad size or null output");
  HIPCHK(launch_normals(seed, stream, n, d_out, (hipStream_t)hstream));
  return 0;
}

int dasp_cent_hc_to_spectrum(const float *pattern_hc, int C, float *spectrum) {
  if (!pattern_hc || !spectrum || C <= 0) return fail(DASP_CENT_EINVAL, "bad arguments");
  std::vector<cd> k((size_t)C * C);
  host_hc_to_spectrum(pattern_hc, C, C, k.data());
  for (size_t q = 0; q < k.size(); ++q) {
    spectrum[2 * q] = (float)k[q].real();
    spectrum[2 * q + 1] = (float)k[q].imag();
  }
  return 0;
}

int dasp_cent_initialise(const dasp_cent_params *pp, dasp_cent_t **out) {
  if (!pp || !out) return fail(DASP_CENT_EINVAL, "null argument");
  *out = nullptr;
  const dasp_cent_params &p = *pp;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(DASP_CENT_ENODEV, "no HIP device");
  {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
        std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      return fail(DASP_CENT_ENODEV, "device %s is not gfx950", prop.gcnArchName);
  }
  dasp_cent *c = new dasp_cent();
  auto bail = [&](int rc) {
    dasp_cent_free(c);
    return rc;
  };
  {   // the fixed scratch array of dasp_cent_run_device_noalloc on this device (looked up once)
    const hipError_t e = noalloc_scratch(&c->noallocScratch);
    if (e != hipSuccess) return bail(fail(DASP_CENT_EHIP, "scratch array address: %s", hipGetErrorString(e)));
  }
  c->p = p;
  c->nsubaps = p.nsubaps;
  c->nsubx = (int)std::sqrt((float)p.nsubaps);
  if (c->nsubx * c->nsubx != p.nsubaps) return bail(fail(DASP_CENT_EINVAL, "nsubaps must be a square"));
  c->ncen = p.ncen;
  c->F = p.fftsize;
  c->clip = p.clipsize;
  c->nimg = p.nimg;
  c->n = p.phasesize;
  c->npup = c->nsubx * c->n;
  c->preBinOrig = p.preBinningFactor;
  c->parabolicFit = p.parabolicFit;
  c->gaussianFit = p.gaussianFit;
  c->gaussianMinVal = p.gaussianMinVal;
  c->gaussianReplaceVal = p.gaussianReplaceVal;
  c->readnoise = p.readnoise;
  c->readbg = p.readbg;
  c->threshType = p.threshType;
  c->addPoisson = p.addPoisson;
  c->noiseFloor = p.noiseFloor;
  c->calsource = p.calsource;
  c->pxlPower = p.pxlPower;
  c->nintegrations = p.nintegrations;
  c->nlatency = p.nlatency;
  c->seed = p.seed == 0 ? (unsigned long)time(nullptr) : p.seed;
  c->opticalBinning = p.opticalBinning;
  c->imageOnly = p.imageOnly;
  c->corrThresh = p.corrThresh;
  c->correlationCentroiding = p.correlationCentroiding;
  if (c->F <= 0 || c->n <= 0 || c->clip <= 0 || c->nimg <= 0 || c->ncen <= 0)
    return bail(fail(DASP_CENT_EINVAL, "sizes must be positive"));
  if (!is_pow2(c->F) || c->F < 8 || c->F > 64 || pad_pow2(c->n) > c->F || c->n > 32)
    return bail(fail(DASP_CENT_EUNSUPPORTED, "fftsize %d / phasesize %d: the device path handles power-of-two fftsize 8..64 with phasesize <= min(32, fftsize)", c->F, c->n));
  if (c->clip > c->F && !p.spotpsf.data)
    return bail(fail(DASP_CENT_EINVAL, "clipsize %d > fftsize %d", c->clip, c->F));
  if (c->nimg > c->clip || c->nimg > kMaxSide) return bail(fail(DASP_CENT_EUNSUPPORTED, "nimg %d must be <= clipsize and <= %d", c->nimg, kMaxSide));
  if (c->ncen > c->nimg) return bail(fail(DASP_CENT_EINVAL, "ncen %d > nimg %d", c->ncen, c->nimg));
  int rc = set_spot(c, p.spotpsf);
  if (rc) return bail(rc);
  // outputs and per-subaperture inputs (checkFloatContigArr / checkIntContigArr at :2132-2149)
  if (!check_f32(p.cents) || numel(p.cents) != (long)c->nsubaps * 2) return bail(fail(DASP_CENT_EINVAL, "cents must be float32 (nsubx,nsubx,2)"));
  if (!p.pupfn.data || numel(p.pupfn) != (long)c->nsubaps * c->n * c->n) return bail(fail(DASP_CENT_EINVAL, "pupfn must be float32 (nsubx,nsubx,n,n)"));
  if (!p.subflag.data || numel(p.subflag) != c->nsubaps) return bail(fail(DASP_CENT_EINVAL, "subflag must be int32 (nsubx,nsubx)"));
  if (!p.bimg.data || numel(p.bimg) != (long)c->nsubaps * c->nimg * c->nimg) return bail(fail(DASP_CENT_EINVAL, "bimg must be float32 (nsubx,nsubx,nimg,nimg)"));
  if (!p.fracSubArea.data || numel(p.fracSubArea) != c->nsubaps) return bail(fail(DASP_CENT_EINVAL, "fracSubArea must be float32 (nsubx,nsubx)"));
  c->h_cents = (float *)p.cents.data;
  c->h_bimg = (float *)p.bimg.data;
  c->pupfn.assign((const float *)p.pupfn.data, (const float *)p.pupfn.data + numel(p.pupfn));
  c->frac.assign((const float *)p.fracSubArea.data, (const float *)p.fracSubArea.data + c->nsubaps);
  c->subflag.assign((const int *)p.subflag.data, (const int *)p.subflag.data + c->nsubaps);
  // phase
  if (p.phs.data) {
    if (p.phs.ndim == 2) c->atmosPhaseType = 0;
    else if (p.phs.ndim == 3) c->atmosPhaseType = 1;
    else return bail(fail(DASP_CENT_EINVAL, "centmodule - phase should be 2D or 3D (for phase+amp)"));
    const int o = c->atmosPhaseType;
    if (p.phs.dims[o] != p.phs.dims[o + 1]) return bail(fail(DASP_CENT_EINVAL, "centmodule - phase should be square"));
    c->inputPhaseSize = (int)p.phs.dims[o];
    if (c->inputPhaseSize != c->npup) {
      // interpolatePhase (:1628-1758): resampled onto the nsubx*phasesize grid on the device
      if (o == 1)
        return bail(fail(DASP_CENT_EUNSUPPORTED, "phase+amplitude on a grid other than nsubx*phasesize (%d) is not supported", c->npup));
      if (c->inputPhaseSize < 3) return bail(fail(DASP_CENT_EINVAL, "phase grid too small to interpolate"));
      c->interpolate = 1;
    }
    if (o == 1 && p.phs.dims[0] != 2) return bail(fail(DASP_CENT_EINVAL, "phase+amplitude should be (2,npup,npup)"));
    c->h_phs = (float *)p.phs.data;
    c->phsElems = numel(p.phs);
  }
  // sig
  c->sig = p.sig;
  if (p.sigArr.data) {
    if (p.sigArr.ndim != 1 || p.sigArr.dims[0] != c->nsubaps) return bail(fail(DASP_CENT_EINVAL, "centmodule - sig arr dimensions must be 1D (nsubx^2)"));
    c->sigArr.assign((const float *)p.sigArr.data, (const float *)p.sigArr.data + c->nsubaps);
    c->haveSigArr = true;
  }
  // sky
  c->skybrightness = p.skybrightness;
  if (p.skybrightnessArr.data) {
    const auto &a = p.skybrightnessArr;
    if (a.ndim != 3 || a.dims[0] != c->nsubaps || a.dims[1] != c->nimg || a.dims[2] != c->nimg)
      return bail(fail(DASP_CENT_EINVAL, "centmodule - skybrightness should be 3D array shape nsubaps,nimg,nimg"));
    c->skybrightness = 0.f;
    if (upload(c->d_skyarr, a.data, numel(a) * 4)) return bail(fail(DASP_CENT_EHIP, "upload failed"));
    c->haveSkyArr = true;
  }
  // correlation
  if (p.corrPattern.data) {
    c->corrPatternSize = (int)p.corrPattern.dims[2];
    rc = set_corr_pattern(c, p.corrPattern);
    if (rc) return bail(rc);
  }
  if (p.corrimg.data) c->h_corrimg = (float *)p.corrimg.data;
  c->corrsize = c->correlationCentroiding ? c->corrPatternSize : c->nimg;
  if (c->correlationCentroiding == 1 && c->haveCorr && !is_pow2(c->corrsize))
    return bail(fail(DASP_CENT_EUNSUPPORTED, "FFT correlation on the device needs a power-of-two pattern size (got %d)", c->corrsize));
  // useBrightest
  c->useBrightest = p.useBrightest;
  if (p.useBrightestArr.data) {
    if (p.useBrightestArr.ndim != 1 || p.useBrightestArr.dims[0] != c->nsubaps)
      return bail(fail(DASP_CENT_EINVAL, "centmodule - useBrightest array should be 1D of shape nsubaps (nsubx^2)"));
    c->useBrightest = 0;
    if (upload(c->d_ubArr, p.useBrightestArr.data, c->nsubaps * 4)) return bail(fail(DASP_CENT_EHIP, "upload failed"));
    c->haveUBArr = true;
  }
  rc = set_cent_weight(c, p.centWeight);
  if (rc) return bail(rc);
  // image input (:3118-3157)
  if (p.inputImage.data) {
    const auto &a = p.inputImage;
    if (a.ndim == 2) {
      if (a.dims[0] * a.dims[1] != (long)c->nimg * c->nimg * c->nsubaps)
        return bail(fail(DASP_CENT_EINVAL, "centmodule - inputImage should be %d x %d", c->nimg * c->nsubx, c->nimg * c->nsubx));
      c->inputDims = 2;
      c->inputNpxlx = (int)a.dims[1];
    } else if (a.ndim == 4) {
      if (a.dims[0] * a.dims[1] != c->nsubaps || a.dims[2] != c->nimg || a.dims[3] != c->nimg)
        return bail(fail(DASP_CENT_EINVAL, "centmodule - inputImage should be (nsubx,nsubx,nimg,nimg)"));
      c->inputDims = 4;
      c->inputNpxlx = (int)a.dims[3];
    } else {
      return bail(fail(DASP_CENT_EINVAL, "inputImage must be 2-D or 4-D"));
    }
    c->h_input = (const float *)a.data;
  }
  if (p.subapLocation.data) {
    if (numel(p.subapLocation) != (long)c->nsubaps * 4)
      return bail(fail(DASP_CENT_EINVAL, "centmodule - subapLocation array should be of size %d", c->nsubaps * 4));
    // the device copies each window into an nimg x nimg image: keep every window inside both
    const int *loc = (const int *)p.subapLocation.data;
    const long rows = c->inputDims == 2 ? p.inputImage.dims[0] : 0, cols = c->inputNpxlx;
    for (int s = 0; s < c->nsubaps && c->inputDims == 2; ++s) {
      const int r0 = loc[4 * s], r1 = loc[4 * s + 1], c0 = loc[4 * s + 2], c1 = loc[4 * s + 3];
      if (r0 < 0 || c0 < 0 || r1 < r0 || c1 < c0 || r1 > rows || c1 > cols || r1 - r0 > c->nimg || c1 - c0 > c->nimg)
        return bail(fail(DASP_CENT_EINVAL, "subapLocation of subaperture %d (%d,%d,%d,%d) is outside the image or larger than nimg", s, r0, r1, c0, c1));
    }
    if (upload(c->d_subapLoc, p.subapLocation.data, (size_t)c->nsubaps * 16)) return bail(fail(DASP_CENT_EHIP, "upload failed"));
  }
  // tilt function (:3074-3081): tmp is a float in the reference
  c->tilt.resize((size_t)c->n * c->n);
  const float tmp = (float)(M_PI / c->F * (c->F + 1.));
  for (int i = 0; i < c->n; ++i)
    for (int j = 0; j < c->n; ++j) c->tilt[i * c->n + j] = tmp * (float)(i + j + 1 - c->n);
  for (int s = 0; s < c->nsubaps; ++s) (c->subflag[s] == 1 ? c->active : c->inactive).push_back(s);
  std::vector<int> winoff(c->active.size());   // phase-window offsets, read by the DFT-GEMM kernels
  for (size_t k = 0; k < c->active.size(); ++k) {
    const int s = c->active[k], sy = s / c->nsubx, sx = s % c->nsubx;
    winoff[k] = sy * c->n * c->npup + sx * c->n;
  }
  compute_totsig(c);
  const int M2 = c->nimg * c->nimg;
  if (upload(c->d_pup, c->pupfn.data(), c->pupfn.size() * 4) || upload(c->d_tilt, c->tilt.data(), c->tilt.size() * 4) ||
      upload(c->d_active, c->active.data(), c->active.size() * 4) ||
      upload(c->d_winoff, winoff.data(), winoff.size() * 4) ||
      upload(c->d_inactive, c->inactive.data(), c->inactive.size() * 4) ||
      c->d_cents.ensure((size_t)c->nsubaps * 2 * 4) || upload(c->d_bimg, c->h_bimg, (size_t)c->nsubaps * M2 * 4))
    return bail(fail(DASP_CENT_EHIP, "device allocation/upload failed"));
  if (c->corrPatternSize > 0 && c->d_corrimg.ensure((size_t)c->nsubaps * c->corrPatternSize * c->corrPatternSize * 4))
    return bail(fail(DASP_CENT_EHIP, "device allocation failed"));
  if (c->h_phs && c->d_phs.ensure(c->phsElems * 4)) return bail(fail(DASP_CENT_EHIP, "device allocation failed"));
  rc = upload_tables(c);
  if (rc) return bail(rc);
  *out = c;
  return 0;
}

int dasp_cent_free(dasp_cent_t *c) {
  if (!c) return 0;
  for (hipEvent_t e : c->evs) (void)hipEventDestroy(e);
  c->evs.clear();
  if (c->corrSide) (void)hipStreamDestroy(c->corrSide);
  if (c->corrFork) (void)hipEventDestroy(c->corrFork);
  if (c->corrJoin) (void)hipEventDestroy(c->corrJoin);
  DevBuf *bufs[] = {&c->d_phs, &c->d_pup, &c->d_tilt, &c->d_active, &c->d_winoff, &c->d_inactive, &c->d_scale, &c->d_totsig,
                    &c->d_sigv, &c->d_skyadd, &c->d_skyarr, &c->d_poisson, &c->d_binA, &c->d_otf, &c->d_corrK, &c->d_corrKp,
                    &c->d_corrPat, &c->d_centWeight, &c->d_refCents, &c->d_calData, &c->d_calBounds,
                    &c->d_calSteps, &c->d_calCoeff, &c->d_ubArr, &c->d_input, &c->d_subapLoc, &c->d_cents,
                    &c->d_bimg, &c->d_corrimg, &c->d_hll, &c->d_zpois, &c->d_zread, &c->d_bscratch, &c->d_cscratch,
                    &c->d_otfk, &c->d_interp};
  for (DevBuf *b : bufs) b->release();
  delete c;
  return 0;
}

int dasp_cent_set_form(dasp_cent_t *c, int form) {
  if (!c) return fail(DASP_CENT_EINVAL, "null handle");
  c->gen++;
  if (form < DASP_CENT_FORM_FFT32 || form > DASP_CENT_FORM_DFT_FP16_1) return fail(DASP_CENT_EINVAL, "unknown form %d", form);
  if (is_dft_form(form) &&
      !(fused_options(c) && !brightest_on(c) && c->spotDim == 0 && dft_supported(c->n, c->F, c->nimg)))
    return fail(DASP_CENT_EUNSUPPORTED, "the DFT-GEMM form covers n=16 with fftsize=clipsize=32 and n=17..32 with "
                                        "fftsize=clipsize=64, nimg=16, the fused options without brightest-pixel selection, "
                                        "no spot kernel, phase only");
  c->form = form;
  return 0;
}

int dasp_cent_set_relative_floor(dasp_cent_t *c, int enable) {
  if (!c) return fail(DASP_CENT_EINVAL, "null handle");
  if (enable && c->threshType != 0) return fail(DASP_CENT_EINVAL, "the relative floor applies to threshType 0 (got %d)", c->threshType);
  c->gen++;
  c->relFloor = enable != 0;
  return 0;
}

int dasp_cent_set_fast_trig(dasp_cent_t *c, int enable) {
  if (!c) return fail(DASP_CENT_EINVAL, "null handle");
  c->gen++;
  c->fastTrig = enable != 0;
  return 0;
}

int dasp_cent_set_grid(dasp_cent_t *c, long image_grid, long corr_grid) {
  if (!c) return fail(DASP_CENT_EINVAL, "null handle");
  c->gen++;   // a step graph holding the old launches is rebuilt
  c->gridImage = image_grid;
  c->gridCorr = corr_grid;
  return 0;
}

int dasp_cent_get_grid(dasp_cent_t *c, long *image_grid, long *corr_grid) {
  if (!c) return fail(DASP_CENT_EINVAL, "null handle");
  if (image_grid) *image_grid = c->gridImage;
  if (corr_grid) *corr_grid = c->gridCorr;
  return 0;
}

int dasp_cent_tune_grid(dasp_cent_t *c, const float *d_phs, long phs_stride, int nframes, float *d_cents,
                        const float *d_zpois, const float *d_zread, int reps, void *stream, double *best_ms) {
  if (!c) return fail(DASP_CENT_EINVAL, "null handle");
  if (reps < 1) return fail(DASP_CENT_EINVAL, "reps must be positive");
  hipStream_t st = (hipStream_t)stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipEventCreate(&e0));
  hipError_t he = hipEventCreate(&e1);
  if (he != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(DASP_CENT_EHIP, "tune: %s", hipGetErrorString(he));
  }
  const bool timing = c->timing;
  c->timing = false;
  // one untimed run (scratch allocation, first-call set-up), then reps timed runs
  auto measure = [&](double &ms) -> int {
    int rc = dasp_cent_run_device(c, d_phs, phs_stride, nframes, d_cents, nullptr, nullptr, d_zpois, d_zread, st);
    if (rc) return rc;
    HIPCHK(hipEventRecord(e0, st));
    for (int r = 0; r < reps && !rc; ++r)
      rc = dasp_cent_run_device(c, d_phs, phs_stride, nframes, d_cents, nullptr, nullptr, d_zpois, d_zread, st);
    if (rc) return rc;
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, e0, e1));
    ms = f / reps;
    return 0;
  };
  static const long kCand[] = {0, 256, 512, 1024, 2048, 4096, -1};
  double best = 0.0;
  int rc = 0;
  for (int stage = 0; stage < 2 && !rc; ++stage) {
    if (stage == 1 && !corr_stage(c)) break;
    long &g = stage == 0 ? c->gridImage : c->gridCorr;
    long keep = g;
    double kept = -1.0;
    for (long cand : kCand) {
      g = cand;
      double ms = 0.0;
      rc = measure(ms);
      if (rc) break;
      if (kept < 0.0 || ms < kept) {
        kept = ms;
        keep = cand;
      }
    }
    g = keep;
    best = kept;
  }
  c->timing = timing;
  c->gen++;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (best_ms) *best_ms = best;
  return 0;
}

int dasp_cent_info(dasp_cent_t *c, int *nactive, int *npup, int *corrsize) {
  if (!c) return fail(DASP_CENT_EINVAL, "null handle");
  if (nactive) *nactive = (int)c->active.size();
  if (npup) *npup = c->npup;
  if (corrsize) *corrsize = c->corrsize;
  return 0;
}

int dasp_cent_set_timing(dasp_cent_t *c, int enable) {
  if (!c) return fail(DASP_CENT_EINVAL, "null handle");
  c->timing = enable != 0;
  return 0;
}

int dasp_cent_kernel_time(dasp_cent_t *c, int reset, double *ms, long *launches) {
  if (!c) return fail(DASP_CENT_EINVAL, "null handle");
  int rc = harvest_events(c);
  if (rc) return rc;
  if (ms) *ms = c->kernel_ms;
  if (launches) *launches = c->kernel_launches;
  if (reset) {
    c->kernel_ms = 0.0;
    c->kernel_launches = 0;
  }
  return 0;
}

int dasp_cent_run(dasp_cent_t *c, float *seconds) {
  if (!c) return fail(DASP_CENT_EINVAL, "null handle");
  const double t0 = now_s();
  // integration cycle (:1774-1788): only the readout step produces images and slopes
  int additive;
  if (c->curInteg == 0) additive = (c->nintegrations + c->nlatency == 1) ? -1 : 0;
  else if (c->curInteg + 1 < c->nintegrations) additive = 1;
  else if (c->curInteg + 1 < c->nintegrations + c->nlatency) additive = 3;
  else additive = 2;
  if (additive == 2 || additive == -1) {
    if (!c->h_phs) return fail(DASP_CENT_EINVAL, "no phase array (image-input module: use runSlope)");
    HIPCHK(hipMemcpy(c->d_phs.p, c->h_phs, c->phsElems * 4, hipMemcpyHostToDevice));
    const bool pois = need_pois(c), rd = need_read(c);
    int rc = draw_noise(c, pois, rd);
    if (rc) return rc;
    Scratch sc;
    rc = device_scratch(c, 1, c->d_bimg.as<float>(), c->d_corrimg.as<float>(), sc);
    if (rc) return rc;
    rc = launch_phase(c, c->d_phs.as<float>(), 0, 1, c->d_cents.as<float>(), sc.bimg, sc.corr,
                      pois ? c->d_zpois.as<float>() : nullptr, rd ? c->d_zread.as<float>() : nullptr, sc, nullptr);
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    rc = copy_back(c);
    if (rc) return rc;
    c->nreadouts++;
  }
  c->curInteg++;
  if (c->curInteg >= c->nintegrations + c->nlatency) c->curInteg = 0;
  if (seconds) *seconds = (float)(now_s() - t0);
  return 0;
}

int dasp_cent_run_slope(dasp_cent_t *c, float *seconds) {
  if (!c) return fail(DASP_CENT_EINVAL, "null handle");
  if (!c->h_input) return fail(DASP_CENT_EINVAL, "centmodule - inputImage is NULL");
  const double t0 = now_s();
  int rc = upload_tables(c);
  if (rc) return rc;
  const int M2 = c->nimg * c->nimg;
  const long nin = (long)c->nsubaps * M2;
  if (upload(c->d_input, c->h_input, nin * 4)) return fail(DASP_CENT_EHIP, "input image upload failed");
  // pixels outside a subapLocation window keep their previous value (:2016-2021)
  HIPCHK(hipMemcpy(c->d_bimg.p, c->h_bimg, nin * 4, hipMemcpyHostToDevice));
  const bool rd = need_read(c) && c->imageOnly && !c->inactive.empty();
  rc = draw_noise(c, false, rd);
  if (rc) return rc;
  BackArgs b;
  fill_back(c, b, 1);
  b.hll = nullptr;
  b.inputImage = c->d_input.as<float>();
  b.inputImageDims = c->inputDims;
  b.inputImagenpxlx = c->inputNpxlx;
  b.subapLocation = c->p.subapLocation.data ? c->d_subapLoc.as<int>() : nullptr;
  b.cents = c->d_cents.as<float>();
  b.bimg = c->d_bimg.as<float>();
  b.corrimg = c->d_corrimg.as<float>();
  HIPCHK(launch_backend(b, nullptr));
  if (!c->inactive.empty()) {
    InactiveArgs ia{};
    ia.inactive = c->d_inactive.as<int>();
    ia.ninactive = (int)c->inactive.size();
    ia.nframes = 1;
    ia.nsubaps = c->nsubaps;
    ia.nimg = c->nimg;
    ia.imageOnly = c->imageOnly;
    ia.readnoise = c->calsource == 0 ? c->readnoise : 0.f;
    ia.readbg = c->readbg;
    ia.floor = c->noiseFloor;
    ia.threshType = kthresh(c);
    ia.zread = rd ? c->d_zread.as<float>() : nullptr;
    ia.noise_stride = nin;
    ia.cents = c->d_cents.as<float>();
    ia.cents_stride = (long)c->nsubaps * 2;
    ia.bimg = c->d_bimg.as<float>();
    ia.bimg_stride = nin;
    HIPCHK(launch_inactive(ia, nullptr));
  }
  HIPCHK(hipDeviceSynchronize());
  rc = copy_back(c);
  if (rc) return rc;
  if (rd) c->nreadouts++;
  if (seconds) *seconds = (float)(now_s() - t0);
  return 0;
}

int dasp_cent_run_device(dasp_cent_t *c, const float *d_phs, long phs_stride, int nframes, float *d_cents,
                         float *d_bimg, float *d_corrimg, const float *d_zpois, const float *d_zread,
                         void *stream) {
  int rc = check_device_run(c, d_phs, phs_stride, nframes, d_cents, d_zpois, d_zread);
  if (rc) return rc;
  Scratch sc;
  rc = device_scratch(c, nframes, d_bimg, d_corrimg, sc);
  if (rc) return rc;
  return launch_phase(c, d_phs, phs_stride, nframes, d_cents, sc.bimg, sc.corr, d_zpois, d_zread, sc, (hipStream_t)stream);
}

int dasp_cent_run_device_noalloc(dasp_cent_t *c, const float *d_phs, long phs_stride, int nframes, float *d_cents,
                                 float *d_bimg, float *d_corrimg, const float *d_zpois, const float *d_zread,
                                 void *stream) {
  int rc = check_device_run(c, d_phs, phs_stride, nframes, d_cents, d_zpois, d_zread);
  if (rc) return rc;
  Scratch sc;
  rc = fixed_scratch(c, nframes, d_bimg, d_corrimg, sc);
  if (rc) return rc;
  // no device allocation below: a handle buffer that would have to grow fails instead, and no
  // timing events are created
  t_forbid_alloc = true;
  t_alloc_refused = false;
  const bool timing = c->timing;
  c->timing = false;
  rc = launch_phase(c, d_phs, phs_stride, nframes, d_cents, sc.bimg, sc.corr, d_zpois, d_zread, sc, (hipStream_t)stream);
  c->timing = timing;
  t_forbid_alloc = false;
  if (rc && t_alloc_refused)
    return fail(rc, "run_device_noalloc: a device buffer of the handle would have to grow (%s)", g_err.c_str());
  return rc;
}

int dasp_cent_update_int(dasp_cent_t *c, int code, long v) {
  if (!c) return fail(DASP_CENT_EINVAL, "null handle");
  c->gen++;
  switch (code) {
    case DASP_CENT_CALSOURCE: c->calsource = (int)v; break;
    case DASP_CENT_ADDPOISSON: c->addPoisson = (int)v; break;
    case DASP_CENT_NCEN:
      if (v < 1 || v > c->nimg) return fail(DASP_CENT_EINVAL, "ncen %ld out of range", v);
      c->ncen = (int)v;
      break;
    case DASP_CENT_NTHREADS: break;   // no host pool
    case DASP_CENT_OPTICALBINNING: c->opticalBinning = (int)v; break;
    case DASP_CENT_CORRELATIONCENTROIDING:
      c->correlationCentroiding = (int)v;
      c->corrsize = v ? c->corrPatternSize : c->nimg;
      if (c->haveCentWeight && c->centWeightSize != c->corrsize) {
        std::fprintf(stderr, "WARNING: centWeightSize INCORRECT: Ignoring centWeight until updated.\n");
        c->haveCentWeight = false;
      }
      break;
    case DASP_CENT_USEBRIGHTEST:
      c->useBrightest = (int)v;
      c->haveUBArr = false;
      break;
    case DASP_CENT_INTEGSTEPS: c->nintegrations = (int)v; break;
    case DASP_CENT_READNOISE: c->readnoise = (float)v; break;
    case DASP_CENT_READBG: c->readbg = (float)v; break;
    case DASP_CENT_NOISEFLOOR: c->noiseFloor = (float)v; break;
    case DASP_CENT_PXLPOWER: c->pxlPower = (float)v; break;
    case DASP_CENT_CORRTHRESH: c->corrThresh = (float)v; break;
    case DASP_CENT_SIG: return dasp_cent_update_float(c, code, (double)v);
    case DASP_CENT_SKYBRIGHTNESS: return dasp_cent_update_float(c, code, (double)v);
    default: return fail(DASP_CENT_EINVAL, "centmodule: Code %d not recognised, nothing updated", code);
  }
  c->tablesDirty = true;
  return 0;
}

int dasp_cent_update_float(dasp_cent_t *c, int code, double v) {
  if (!c) return fail(DASP_CENT_EINVAL, "null handle");
  c->gen++;
  switch (code) {
    case DASP_CENT_SIG:
      c->sig = (float)v;   // a sig array given earlier stays in use, as in py_update (:2596-2606)
      compute_totsig(c);
      break;
    case DASP_CENT_READNOISE: c->readnoise = (float)v; break;
    case DASP_CENT_READBG: c->readbg = (float)v; break;
    case DASP_CENT_NOISEFLOOR: c->noiseFloor = (float)v; break;
    case DASP_CENT_SKYBRIGHTNESS:
      c->skybrightness = (float)v;
      c->haveSkyArr = false;
      break;
    case DASP_CENT_PXLPOWER: c->pxlPower = (float)v; break;
    case DASP_CENT_CORRTHRESH: c->corrThresh = (float)v; break;
    default: return dasp_cent_update_int(c, code, (long)v);
  }
  c->tablesDirty = true;
  return 0;
}

int dasp_cent_update_array(dasp_cent_t *c, int code, const dasp_cent_array *va) {
  if (!c || !va) return fail(DASP_CENT_EINVAL, "null argument");
  c->gen++;
  const dasp_cent_array &a = *va;
  int rc = 0;
  switch (code) {
    case DASP_CENT_SIG:
      if (!a.data || a.ndim != 1 || a.dims[0] != c->nsubaps) return fail(DASP_CENT_EINVAL, "sigArr should be 1D, size %d", c->nsubaps);
      c->sigArr.assign((const float *)a.data, (const float *)a.data + c->nsubaps);
      c->haveSigArr = true;
      break;
    case DASP_CENT_SKYBRIGHTNESS:
      if (!a.data || a.ndim != 3 || a.dims[0] != c->nsubaps || a.dims[1] != c->nimg || a.dims[2] != c->nimg)
        return fail(DASP_CENT_EINVAL, "Error - centmodule - skybrightness should be 3D array shape nsubaps,nimg,nimg");
      if (upload(c->d_skyarr, a.data, numel(a) * 4)) return fail(DASP_CENT_EHIP, "upload failed");
      c->skybrightness = 0.f;
      c->haveSkyArr = true;
      break;
    case DASP_CENT_SPOTPSF:
      rc = set_spot(c, a);
      if (rc) return rc;
      compute_totsig(c);
      break;
    case DASP_CENT_CENTWEIGHT:
      rc = set_cent_weight(c, a);
      if (rc) return rc;
      break;
    case DASP_CENT_CORRPATTERN:
      if (!a.data) {
        c->haveCorr = false;
        break;
      }
      if (a.ndim != 4 || c->corrPatternSize != a.dims[2])
        return fail(DASP_CENT_EINVAL, "Error - centmodule - corrPattern has changed size - please don't do that (or recode)!");
      rc = set_corr_pattern(c, a);
      if (rc) return rc;
      break;
    case DASP_CENT_REFCENTS:
      if (!a.data) {
        c->haveRef = false;
        break;
      }
      if (numel(a) < (long)c->nsubaps * 2) return fail(DASP_CENT_EINVAL, "refCents must hold nsubaps*2 floats");
      if (upload(c->d_refCents, a.data, (size_t)c->nsubaps * 8)) return fail(DASP_CENT_EHIP, "upload failed");
      c->haveRef = true;
      break;
    case DASP_CENT_CALCOEFF:
      if (!a.data) {
        c->haveCalCoeff = false;
        c->calNCoeff = 0;
        break;
      }
      if (a.ndim != 4 || a.dims[2] != 2 || a.dims[0] * a.dims[1] != c->nsubaps)
        return fail(DASP_CENT_EINVAL, "Error: centmodule - calCoeff should be 4d (nsubx,nsubx,2,ncoeff)");
      c->calNCoeff = (int)a.dims[3];
      if (upload(c->d_calCoeff, a.data, numel(a) * 4)) return fail(DASP_CENT_EHIP, "upload failed");
      c->haveCalCoeff = true;
      break;
    case DASP_CENT_USEBRIGHTEST:
      if (!a.data || a.ndim != 1 || a.dims[0] != c->nsubaps)
        return fail(DASP_CENT_EINVAL, "Error: centmodule - useBrightest array should be 1D of shape nsubaps (nsubx^2)");
      if (upload(c->d_ubArr, a.data, (size_t)c->nsubaps * 4)) return fail(DASP_CENT_EHIP, "upload failed");
      c->useBrightest = 0;
      c->haveUBArr = true;
      break;
    case DASP_CENT_OUTPUTPHASEARR:   // (:2832-2848) the interpolated phase, for verification
      c->h_outphase = nullptr;
      if (!a.data) break;
      if (a.ndim != 2 || a.dims[0] != c->npup || a.dims[1] != c->npup)
        return fail(DASP_CENT_EINVAL, "outputPhaseArr should be 2D of size %d x %d", c->npup, c->npup);
      c->h_outphase = (float *)a.data;
      break;
    default: return fail(DASP_CENT_EINVAL, "centmodule: Code %d not recognised, nothing updated", code);
  }
  c->tablesDirty = true;
  return 0;
}

int dasp_cent_update_caldata(dasp_cent_t *c, const dasp_cent_array *cd_, const dasp_cent_array *cb,
                             const dasp_cent_array *cs) {
  if (!c) return fail(DASP_CENT_EINVAL, "null handle");
  c->gen++;
  if (!cd_ || !cd_->data) {
    c->haveCalData = false;
    c->calNSteps = 0;
    return 0;
  }
  if (!cb || !cs || !cb->data || !cs->data) return fail(DASP_CENT_EINVAL, "Error: centmodule - calibration stuff not a tuple");
  if (cs->ndim != 1) return fail(DASP_CENT_EINVAL, "Error: centmodule - calStep should be 1d");
  const int steps = (int)cs->dims[0];
  if (cd_->ndim != 4 || cd_->dims[0] != 2 || cd_->dims[1] * cd_->dims[2] != c->nsubaps || cd_->dims[3] != steps)
    return fail(DASP_CENT_EINVAL, "Error: centmodule - calData should be 4d, (2,nsubx,nsubx,steps)");
  if (cb->ndim != 4 || cb->dims[0] != 2 || cb->dims[1] * cb->dims[2] != c->nsubaps || cb->dims[3] != 2)
    return fail(DASP_CENT_EINVAL, "Error: centmodule - calBounds should be 4d (2,nsubx,nsubx,2)");
  const int *bnd = (const int *)cb->data;
  for (long q = 0; q < 2L * c->nsubaps; ++q)
    if (bnd[2 * q] < 0 || bnd[2 * q + 1] < bnd[2 * q] || bnd[2 * q + 1] >= steps)
      return fail(DASP_CENT_EINVAL, "calBounds entry %ld (%d,%d) outside [0,%d)", q, bnd[2 * q], bnd[2 * q + 1], steps);
  if (upload(c->d_calData, cd_->data, numel(*cd_) * 4) || upload(c->d_calBounds, cb->data, numel(*cb) * 4) ||
      upload(c->d_calSteps, cs->data, numel(*cs) * 4))
    return fail(DASP_CENT_EHIP, "upload failed");
  c->calNSteps = steps;
  c->haveCalData = true;
  return 0;
}

int dasp_cent_testcorrelation(const float *img, float *corr, int n) {
  // py_testcorrelation (:3275-3309) on the host, double precision: corr <- HC correlation
  if (!img || !corr || n <= 0) return fail(DASP_CENT_EINVAL, "bad arguments");
  std::vector<cd> K((size_t)n * n), X((size_t)n * n);
  host_hc_to_spectrum(corr, n, n, K.data());
  for (size_t q = 0; q < X.size(); ++q) X[q] = cd(img[q], 0);
  host_dft2(X.data(), n, n, -1);
  for (size_t q = 0; q < X.size(); ++q) X[q] *= K[q];
  host_dft2(X.data(), n, n, +1);
  for (size_t q = 0; q < X.size(); ++q) corr[q] = (float)X[q].real();
  return 0;
}


int dasp_cent_batch_create(const dasp_cent_batch_item *items, int nitems, dasp_cent_batch_t **out) {
  if (!out) return fail(DASP_CENT_EINVAL, "null output pointer");
  *out = nullptr;
  if (!items || nitems < 1) return fail(DASP_CENT_EINVAL, "a batch needs at least one item");
  for (int i = 0; i < nitems; ++i) {
    if (!items[i].handle) return fail(DASP_CENT_EINVAL, "batch item %d: null handle", i);
    for (int j = 0; j < i; ++j)
      if (items[j].handle == items[i].handle) return fail(DASP_CENT_EINVAL, "batch items %d and %d share a handle", j, i);
  }
  dasp_cent_batch *b = new dasp_cent_batch;
  b->items.assign(items, items + nitems);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
  for (int k = 0; k < dasp_cent_batch::kSide && e == hipSuccess; ++k) {
    e = hipStreamCreateWithFlags(&b->side[k], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  if (e != hipSuccess) {
    dasp_cent_batch_free(b);
    return fail(DASP_CENT_EHIP, "batch set-up: %s", hipGetErrorString(e));
  }
  int rc = batch_layout(b);
  if (!rc) rc = batch_capture(b);
  if (rc) {
    std::string msg = g_err;
    dasp_cent_batch_free(b);
    return fail(rc, "%s", msg.c_str());
  }
  *out = b;
  return 0;
}

int dasp_cent_batch_tune(dasp_cent_batch_t *b, void *stream, int reps) {
  if (!b) return fail(DASP_CENT_EINVAL, "null batch");
  if (reps < 1) reps = 1;
  int rc = batch_refresh(b);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  rc = 0;
  for (BatchLaunch &L : b->launches) {
    float best = -1.f;
    int bw = L.waves, bc = L.chunk;
    long bg = L.wgs;
    struct Cand {
      int waves, chunk;
      long wgs;
    };
    std::vector<Cand> cand;
    for (int waves : {1, 2, 4}) {
      const int per_cu = batch_resident_wgs(L.np, L.F, L.nimg, L.spot, waves, L.fast);
      const long full = std::min((L.groups + waves - 1) / waves, 0x7fffffffL);
      for (int k = 1; k <= per_cu; ++k) cand.push_back({waves, 0, std::min((long)k * b->cus, full)});
      cand.push_back({waves, 0, full});   // one group per wave: no persistence
      for (int chunk : {2, 8}) cand.push_back({waves, chunk, std::min((long)std::max(per_cu, 1) * b->cus, full)});
    }
    for (const Cand &cd : cand) {
      hipError_t e = batch_launch(L, (int)L.items.size(), cd.waves, cd.chunk, cd.wgs, st);   // warm-up
      if (e == hipSuccess) e = hipEventRecord(e0, st);
      for (int r = 0; r < reps && e == hipSuccess; ++r) e = batch_launch(L, (int)L.items.size(), cd.waves, cd.chunk, cd.wgs, st);
      if (e == hipSuccess) e = hipEventRecord(e1, st);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) {
        rc = fail(DASP_CENT_EHIP, "tuning launch failed: %s", hipGetErrorString(e));
        break;
      }
      ms /= reps;
      if (best < 0.f || ms < best) {
        best = ms;
        bw = cd.waves;
        bc = cd.chunk;
        bg = cd.wgs;
      }
    }
    if (rc) break;
    L.waves = bw;
    L.chunk = bc;
    L.wgs = bg;
    L.ms = best;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  return batch_capture(b);
}

int dasp_cent_batch_run(dasp_cent_batch_t *b, void *stream) {
  if (!b) return fail(DASP_CENT_EINVAL, "null batch");
  int rc = batch_refresh(b);
  if (rc) return rc;
  HIPCHK(hipGraphLaunch(b->exec, (hipStream_t)stream));
  return 0;
}

int dasp_cent_batch_launches(dasp_cent_batch_t *b) {
  if (!b) return fail(DASP_CENT_EINVAL, "null batch");
  return (int)b->launches.size();
}

int dasp_cent_batch_set_config(dasp_cent_batch_t *b, int i, int waves_per_wg, long workgroups, int chunk) {
  if (!b) return fail(DASP_CENT_EINVAL, "null batch");
  int rc = batch_refresh(b);
  if (rc) return rc;
  if (i < 0 || i >= (int)b->launches.size()) return fail(DASP_CENT_EINVAL, "launch %d out of range", i);
  if (waves_per_wg != 1 && waves_per_wg != 2 && waves_per_wg != 4)
    return fail(DASP_CENT_EINVAL, "waves per workgroup must be 1, 2 or 4");
  if (workgroups < 1 || workgroups > 0x7fffffffL || chunk < 0) return fail(DASP_CENT_EINVAL, "bad workgroups/chunk");
  BatchLaunch &L = b->launches[i];
  L.waves = waves_per_wg;
  L.wgs = workgroups;
  L.chunk = chunk;
  return batch_capture(b);
}

int dasp_cent_batch_launch_info(dasp_cent_batch_t *b, int i, int shape[5], int *nwfs, long *groups,
                                int *waves_per_wg, long *workgroups, int *chunk, float *ms) {
  if (!b) return fail(DASP_CENT_EINVAL, "null batch");
  if (i < 0 || i >= (int)b->launches.size()) return fail(DASP_CENT_EINVAL, "launch %d out of range", i);
  const BatchLaunch &L = b->launches[i];
  if (shape) {
    shape[0] = L.np;
    shape[1] = L.F;
    shape[2] = L.nimg;
    shape[3] = L.spot;
    shape[4] = L.fast;
  }
  if (nwfs) *nwfs = (int)L.items.size();
  if (groups) *groups = L.groups;
  if (waves_per_wg) *waves_per_wg = L.waves;
  if (workgroups) *workgroups = L.wgs;
  if (chunk) *chunk = L.chunk;
  if (ms) *ms = L.ms;
  return 0;
}

int dasp_cent_batch_others(dasp_cent_batch_t *b) {
  if (!b) return fail(DASP_CENT_EINVAL, "null batch");
  return (int)b->others.size();
}

int dasp_cent_batch_free(dasp_cent_batch_t *b) {
  if (!b) return 0;
  batch_drop_graph(b);
  for (BatchLaunch &L : b->launches) {
    L.desc.release();
    L.counters.release();
  }
  b->inactive.release();
  if (b->cap) (void)hipStreamDestroy(b->cap);
  for (int k = 0; k < dasp_cent_batch::kSide; ++k) {
    if (b->side[k]) (void)hipStreamDestroy(b->side[k]);
    if (b->join[k]) (void)hipEventDestroy(b->join[k]);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  delete b;
  return 0;
}

}  // extern "C"
