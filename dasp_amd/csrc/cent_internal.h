// cent_internal.h - argument blocks shared by the host engine (cent_api.cpp) and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dasp {

constexpr int kMaxSide = 64;  // largest psfsize/clipsize/nimg/corrsize handled on the device

// Size of the fixed per-device scratch array of the allocation-free device run
// (dasp_cent_run_device_noalloc, cent_noalloc.hip).
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
// Its address on the current device (looked up once per device).
hipError_t noalloc_scratch(void **out);

// Read of a read-only input at a wave-uniform index through the scalar cache (constant address
// space). Such reads are counted by the scalar-memory counter, so waiting for them never waits
// for the wave's outstanding vector loads (prefetches of the next item).
template <typename T>
__device__ __forceinline__ T scalar_read(const T *p, long i) {
  typedef const __attribute__((address_space(4))) T CT;
  return ((CT *)p)[i];
}

// Global-memory view of a wave-uniform address, rebuilt from its first lane's value. Row bases
// made with it stay in scalar registers, and a load at base + (32-bit lane offset) then takes the
// scalar-base plus vector-offset form; left to itself the compiler chains consecutive row
// addresses through 64-bit vector adds from the first row's per-lane address.
typedef const __attribute__((address_space(1))) char GlobalBytes;
__device__ __forceinline__ GlobalBytes *uniform_global(const void *p) {
  const unsigned long long v = (unsigned long long)p;
  const unsigned long long lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
  const unsigned long long hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
  return (GlobalBytes *)((hi << 32) | lo);
}
// A lane offset that is loop-invariant would be zero-extended once, before the loop, and the
// 64-bit value no longer matches the scalar-base load form; an empty asm in the loop body keeps
// the extension beside the loads. Call once per item on the offsets used by load_at.
__device__ __forceinline__ void pin_offset(unsigned &byte_offset) { asm volatile("" : "+v"(byte_offset)); }
__device__ __forceinline__ float load_at(GlobalBytes *base, unsigned byte_offset) {
  return *(const __attribute__((address_space(1))) float *)(base + byte_offset);
}

// exp(i th) for the phase argument. Accurate path: sincosf. Fast path: th in revolutions reduced
// to |r| <= 1/2 as r = th*c_hi - k + th*c_lo (c_hi + c_lo = 1/(2 pi) to 2^-48, k = rint(th*c_hi);
// the first product is exact inside the fma), then the hardware sine/cosine of r revolutions.
// Reduction error at most 2.6e-7 rad for |th| <= 400.
template <bool FAST>
__device__ __forceinline__ void phase_sincos(float th, float *sn, float *cs) {
  if constexpr (FAST) {
    const float k = rintf(th * 0.15915494309189535f);
    float r = fmaf(th, 0.15915494309189535f, -k);
    r = fmaf(th, 6.4206382e-09f, r);
    *sn = __builtin_amdgcn_sinf(r);
    *cs = __builtin_amdgcn_cosf(r);
  } else {
    sincosf(th, sn, cs);
  }
}

// Image formation: phase -> |FFT|^2 placed as the reference's hll (cmod/src/centmodule.c:473-509).
struct FormArgs {
  const float *phs;      // frame 0 phase (npup x npup)
  const float *amp;      // frame 0 amplitude plane or nullptr
  long phs_stride;       // floats between frames
  const float *pup;      // (nsubaps, n, n)
  const float *tilt;     // (n, n) radians, float as the reference stores it
  const int *active;     // active subaperture indices
  int nactive;
  long nitems;           // nframes * nactive
  int nsubx, n, npup;
  int padded;            // paddedsize: side of hll (clip, or psfsize with a spot kernel)
  float *hll;            // (nitems, padded, padded)
  int fast_trig;         // the phase factor's sin/cos as in FusedArgs::fast_trig
};

// Everything after image formation, one (frame, subaperture) item per workgroup.
struct BackArgs {
  const float *hll;      // (nitems, padded, padded), or nullptr in image-input mode
  const int *active;
  int nactive;
  long nitems;
  int nsubx, nsubaps, n, F, clip, nimg, ncen, psf, padded, pbf, corrsize, npup;
  // LGS spot kernel: full Hermitian spectrum (psf x psf complex) per subap or shared
  const float2 *otf;
  int otf_per_subap;
  // binning operator (nimg x clip); binB > 0 when it is the plain binB x binB block sum
  const float *binA;
  int binB;
  // normalisation
  const float *scale;    // per subap factor sig*frac/totsig (1 where totsig <= 0)
  const float *totsig;   // per subap theoretical total
  const float *sigv;     // per subap sig*frac or sigArr
  const float *pup;      // (nsubaps, n, n): for the phase+amplitude total
  const float *amp;      // frame 0 amplitude plane or nullptr
  long phs_stride;
  // sky
  const float *skyadd;   // per subap sky*frac (0 when <= 0)
  const float *skyarr;   // (nsubaps, nimg, nimg) or nullptr
  // noise (pre-drawn N(0,1), per frame nsubaps*nimg*nimg)
  const unsigned char *poisson_on;  // per subap flag (addPoisson && !calsource && signal)
  const float *zpois;
  const float *zread;
  long noise_stride;
  float readnoise;       // already multiplied by (calsource==0)
  float readbg, floor;
  int threshType;
  int opticalBinning;
  int calsource;
  int useBrightest;
  const int *useBrightestArr;
  float pxlPower;
  int imageOnly;
  // correlation
  int corrMode;
  float corrThresh;
  const float2 *corrK;   // (nsubaps, C, C) spectrum for mode 1
  const float *corrPattern;  // (nsubaps, C, C) raw pattern for modes 2..5
  // centroid
  const float *centWeight;
  int centWeightDim;     // 0, 2 or 4
  int fitsize;           // effective fit size (3.. ) or -1 for the 3x3 quadratic interpolation
  int doFit;             // parabolicFit!=0 || gaussianFit>0
  int gaussianFit;
  float gaussianMinVal, gaussianReplaceVal;
  float fitMx[30];
  const float *calData;  // (2, nsubaps, calNSteps)
  const int *calBounds;  // (2, nsubaps, 2)
  const float *calSteps;
  int calNSteps;
  const float *calCoeff; // (nsubaps, 2, calNCoeff)
  int calNCoeff;
  const float *refCents;
  // image-input mode (centroidsFromImage, cmod/src/centmodule.c:1987-2127)
  const float *inputImage;
  int inputImageDims;    // 0 = phase mode, 2 or 4
  int inputImagenpxlx;
  const int *subapLocation;
  // outputs (frame f adds f*stride)
  float *cents;
  long cents_stride;
  float *bimg;
  long bimg_stride;
  float *corrimg;
  long corr_stride;
};

// Inactive subapertures: zero slopes, or read-noise-only images when imageOnly != 0.
struct InactiveArgs {
  const int *inactive;
  int ninactive;
  int nframes;
  int nsubaps, nimg;
  int imageOnly;
  float readnoise, readbg, floor;
  int threshType;
  const float *zread;
  long noise_stride;
  float *cents;
  long cents_stride;
  float *bimg;
  long bimg_stride;
};

// Fused fast path (no spot kernel, no prebinning, integer binning from a clip == fftsize image,
// plain or weighted CoG, or FFT correlation + parabolic fit).
struct FusedArgs {
  const float *phs;
  long phs_stride;
  const float *pup;      // (nsubaps, n, n)
  const float *tilt;     // (n, n)
  float tiltstep;        // the reference's float tmp = pi/F*(F+1): tilt[i][j] = tmp*(i+j+1-n)
  int fast_trig;         // 1: reduction to revolutions + hardware sin/cos instead of sincosf
  int slow_detector;     // any of noise, sky array, relative floor, power, weights
  const int *active;
  const int *winoff;     // per active index: (sy*n)*npup + sx*n, the phase window's offset in a frame
  int nactive;
  long nitems;
  int nsubx, n, npup, nimg, ncen;
  const float *scale;
  const float *skyadd;
  const float *skyarr;
  const unsigned char *poisson_on;
  const float *zpois;
  const float *zread;
  long noise_stride;
  float readnoise, readbg, floor;
  int threshType;
  int calsource;
  float pxlPower;
  const float *centWeight;
  int centWeightDim;
  const float *refCents;
  float *cents;
  long cents_stride;
  float *bimg;           // may be nullptr (images not returned)
  long bimg_stride;
  // LGS spot kernel (psfsize == clipsize == fftsize == 2 * phase pad): per subaperture
  // (F/2 + 1) rows of F complex, see spot_convolve in fused_item.h
  const float2 *otfk;
  int otf_per_subap;
  int no_cents;          // slopes come from the correlation/fit stage (cent_corr.hip)
  // inactive sub-apertures whose slopes the image kernel zeroes for every frame, or nullptr when
  // the separate inactive kernel does it (multi-WFS step, image-only modes)
  const int *inactive;
  int ninactive;
  // 1 when the detector is the identity up to a positive scale for every sub-aperture (no
  // image returned, nothing added, a threshold that changes nothing, the whole image as CoG
  // window, every scale > 0): the CoG may then be taken on the bins themselves
  int bins_only;
  // the brightest-pixel selection (selectBrightestPixels): a count for every sub-aperture, or one
  // per sub-aperture (useBrightestArr, non-null), 0 for none; negative counts subtract the largest
  // value below the cut. Both off when calsource is set.
  int useBrightest;
  const int *useBrightestArr;
};

// Slopes of the inactive sub-apertures set to zero for every frame (centroidsFromPhase gives
// them cents = 0, cmod/src/centmodule.c:1964-1969), spread over all threads of the grid at the
// start of an image kernel, so a step needs no second launch for them.
__device__ __forceinline__ void zero_inactive_slopes(const FusedArgs &a) {
  if (a.inactive == nullptr) return;
  const long total = (a.nitems / a.nactive) * a.ninactive;
  const long step = (long)gridDim.x * blockDim.x;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += step) {
    const long f = e / a.ninactive;
    float *co = a.cents + f * a.cents_stride + 2 * a.inactive[e - f * a.ninactive];
    co[0] = 0.f;
    co[1] = 0.f;
  }
}

// Correlation / fit stage after the fused kernels (cent_corr.hip).
struct CorrArgs {
  const float *img;      // final detector images (frames, nsubaps, nimg, nimg)
  long img_stride;
  const int *active;
  int nactive;
  long nitems;
  int nimg, ncen;
  const float2 *corrK;   // (nsubaps, N, N) pattern spectrum, mode 1
  const float4 *corrKp;  // (nsubaps, N, N/2) two-image table (alpha, beta) for N = 8, 16 (or null)
  float corrThresh;
  int doFit, fitsize, gaussianFit;
  float gaussianMinVal, gaussianReplaceVal;
  float fitMx[30];
  const float *centWeight;
  int centWeightDim;
  const float *refCents;
  float *cents;
  long cents_stride;
  float *corrimg;
  long corr_stride;
};
// interpolatePhase on the whole output grid (cent_interp.hip).
struct InterpArgs {
  const float *in;       // (m, m) per frame
  long in_stride;
  int m;
  float *out;            // (tot, tot) per frame, contiguous frames
  int tot;               // nsubx * phasesize
  float scale;           // (m - 1) / (tot - 1), float as the reference stores it
  int nframes;
};
hipError_t launch_interp(const InterpArgs &a, hipStream_t st);

// n standard normals of the (seed, stream) noise stream into device memory (cent_noise.hip),
// the host generator's sequence
hipError_t launch_normals(uint64_t seed, uint64_t stream, long n, float *out, hipStream_t st);

bool corr_fit_supported(int n);
hipError_t launch_corr_fit(const CorrArgs &a, int n, bool corr, hipStream_t st, long grid = 0);

// Workgroups of a persistent grid-stride launch: `natural` (one per work unit) capped at `cap`
// (the kernel's default persistent grid); a tuned override g > 0 caps at g instead, g < 0 lifts
// the cap. Rounded up to a whole multiple of the 8 XCDs (workgroup b runs on XCD b % 8, and the
// kernels split their work by XCD, so every XCD gets the same number of workgroups).
inline long launch_grid(long natural, long cap, long g) {
  constexpr long xcds = 8;
  const long c = g > 0 ? g : (g < 0 ? natural : cap);
  long b = natural < c ? natural : c;
  if (b < 1) b = 1;
  return (b + xcds - 1) / xcds * xcds;
}

// Multi-WFS persistent step (batch_kernel.h): one descriptor per WFS in a shape's launch.
struct BatchDesc {
  FusedArgs args;
  long group0;           // first queue entry of this WFS (groups of 64/NP work items)
};

// Kernel argument blocks travel by value between separately compiled objects: a layout change
// must come with a full rebuild (dasp_amd/build.py compiles every source each time).
static_assert(sizeof(FusedArgs) == 288, "FusedArgs layout changed: update this pin and rebuild all objects");
static_assert(sizeof(BatchDesc) == 296, "BatchDesc layout changed: update this pin and rebuild all objects");

struct BatchArgs {
  const BatchDesc *desc; // ascending group0, desc[0].group0 == 0
  int ndesc;
  long total_groups;
  unsigned long long *counters;   // [0] next queue entry, [1] waves finished; zero between launches
};

static_assert(sizeof(BatchArgs) == 32, "BatchArgs layout changed: update this pin and rebuild all objects");

// Inactive subapertures of one WFS in a multi-WFS step: entries [start, start + nframes*ninactive).
struct InactiveDesc {
  const int *inactive;
  int ninactive;
  float *cents;
  long cents_stride;
  long start;
};
hipError_t launch_batch_inactive(const InactiveDesc *d, int nd, long total, hipStream_t st);

bool batch_supported(int n_pad, int F, int nimg, bool spot);
int batch_resident_wgs(int n_pad, int F, int nimg, bool spot, int waves_per_wg, int fast);   // per CU
// chunk: 0 = static schedule, > 0 = entries claimed per atomic (see WaveQueue in batch_kernel.h)
hipError_t launch_batch(const BatchArgs &a, int n_pad, int F, int nimg, bool spot, int waves_per_wg, int chunk,
                        int fast, long workgroups, hipStream_t st);

hipError_t launch_form(const FormArgs &a, int n_pad, int F, hipStream_t st);
hipError_t launch_backend(const BackArgs &a, hipStream_t st);
hipError_t launch_inactive(const InactiveArgs &a, hipStream_t st);
hipError_t launch_fill(float *p, long n, float v, hipStream_t st);
bool fused_supported(int n_pad, int F, int nimg, bool spot);
hipError_t launch_fused(const FusedArgs &a, int n_pad, int F, bool spot, hipStream_t st, long grid = 0);
bool dft_supported(int n, int F, int nimg);
// comp: the three-product (compensated) scheme, else one product of the FP16 roundings (bf16 = 0)
hipError_t launch_dft(const FusedArgs &a, int bf16, bool comp, hipStream_t st, long grid = 0);
// n in 17..32, fftsize 64, nimg 16 (cent_dft32.hip); reached through launch_dft
bool dft32_supported(int n, int F, int nimg);
hipError_t launch_dft32(const FusedArgs &a, int bf16, bool comp, hipStream_t st, long grid = 0);

}  // namespace dasp
