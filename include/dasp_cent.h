/*
 * dasp_cent.h - C ABI of the MI355X (gfx950) Shack-Hartmann image-formation and centroiding
 * engine. It replaces the CPython module cmod.cent of DASP (cmod/src/centmodule.c) behind the
 * util.centcmod.centcmod construct/run/update contract (util/centcmod.py:41-145).
 *
 * Conventions
 *   - Every entry point returns 0 on success and a negative DASP_CENT_E* code on failure; the
 *     message of the last failure on the calling thread is returned by dasp_cent_last_error().
 *   - "host" pointers are borrowed, exactly as the reference borrows numpy buffers
 *     (cmod/src/centmodule.c:2978-3000): the caller keeps them alive until dasp_cent_free() or
 *     until it replaces them through dasp_cent_update_*(). Outputs (cents, bimg, corrimg) are
 *     written in place by dasp_cent_run().
 *   - "device" pointers are HIP device addresses; stream arguments are hipStream_t passed as
 *     void* (NULL = the legacy default stream).
 *   - Arrays are C-contiguous float32 / int32 in the reference layouts.
 */
#ifndef DASP_CENT_H
#define DASP_CENT_H

#ifdef __cplusplus
extern "C" {
#endif

/* Bumped whenever a structure below changes layout; cent.py refuses a library of another version. */
#define DASP_CENT_ABI_VERSION 2

/* error codes */
#define DASP_CENT_OK 0
#define DASP_CENT_EINVAL (-1)      /* bad argument, shape or type */
#define DASP_CENT_EUNSUPPORTED (-2) /* valid for the reference, not supported by this build */
#define DASP_CENT_EHIP (-3)        /* HIP runtime error */
#define DASP_CENT_ENOMEM (-4)
#define DASP_CENT_ENODEV (-5)      /* no gfx950 device */

/* update codes: identical values to cmod/src/centmodule.c:109-130 and util/centcmod.py:19-40 */
#define DASP_CENT_CALSOURCE 1
#define DASP_CENT_SIG 2
#define DASP_CENT_ADDPOISSON 3
#define DASP_CENT_NCEN 4
#define DASP_CENT_READNOISE 5
#define DASP_CENT_READBG 6
#define DASP_CENT_NOISEFLOOR 7
#define DASP_CENT_SKYBRIGHTNESS 8
#define DASP_CENT_PXLPOWER 9
#define DASP_CENT_SPOTPSF 10
#define DASP_CENT_NTHREADS 11
#define DASP_CENT_OPTICALBINNING 12
#define DASP_CENT_CENTWEIGHT 13
#define DASP_CENT_CORRELATIONCENTROIDING 14
#define DASP_CENT_CORRTHRESH 15
#define DASP_CENT_CORRPATTERN 16
#define DASP_CENT_CALDATA 17
#define DASP_CENT_REFCENTS 18
#define DASP_CENT_CALCOEFF 19
#define DASP_CENT_USEBRIGHTEST 20
#define DASP_CENT_INTEGSTEPS 21
#define DASP_CENT_OUTPUTPHASEARR 22

/* compute forms (extension; the reference has only its FFTW path) */
#define DASP_CENT_FORM_FFT32 0     /* fused FP32 FFT in registers/LDS */
#define DASP_CENT_FORM_DFT_BF16 1  /* DFT-matrix GEMMs on MFMA, BF16 hi/lo, FP32 accumulate */
#define DASP_CENT_FORM_DFT_FP16 2  /* DFT-matrix GEMMs on MFMA, FP16 hi/lo, FP32 accumulate */
#define DASP_CENT_FORM_FFT32_GENERAL 3 /* FP32 FFT through the general (unfused) kernels only */
#define DASP_CENT_FORM_DFT_FP16_1 4 /* DFT-matrix GEMMs on MFMA, one product of the FP16 roundings */

typedef struct dasp_cent dasp_cent_t;

/* An n-dimensional float32/int32 array argument; data==NULL stands for Python's None. */
typedef struct {
  void *data;
  int ndim;
  long dims[4];
} dasp_cent_array;

/*
 * The 41 arguments of cent.initialise in the reference order
 * (cmod/src/centmodule.c:2876-2883, format "iiiiiiiffifOOifiilOO!OO!O!O!O!iOifOOiiOiiiffOO";
 * caller util/centcmod.py:115). Arguments that Python passes as "float or array" have both forms:
 * the array form wins when its data pointer is non-NULL.
 */
typedef struct {
  int nthreads;              /* host threads of the reference; ignored by the GPU engine */
  int nsubaps;               /* nsubx*nsubx */
  int ncen, fftsize, clipsize, nimg, phasesize;
  float readnoise, readbg;
  int addPoisson;
  float noiseFloor;
  float sig;                 /* used when sigArr.data == NULL */
  dasp_cent_array sigArr;    /* f32 (nsubaps,) */
  float skybrightness;       /* used when skybrightnessArr.data == NULL */
  dasp_cent_array skybrightnessArr; /* f32 (nsubaps,nimg,nimg) */
  int calsource;
  float pxlPower;
  int nintegrations, nlatency;
  unsigned long seed;
  dasp_cent_array phs;       /* f32 (npup,npup) phase or (2,npup,npup) phase+amplitude; NULL = image input */
  dasp_cent_array pupfn;     /* f32 (nsubx,nsubx,n,n) */
  dasp_cent_array spotpsf;   /* f32 (P,P) or (nsubx,nsubx,P,P), P>=clipsize, or NULL */
  dasp_cent_array cents;     /* f32 (nsubx,nsubx,2) output */
  dasp_cent_array subflag;   /* i32 (nsubx,nsubx) */
  dasp_cent_array bimg;      /* f32 (nsubx,nsubx,nimg,nimg) output */
  dasp_cent_array fracSubArea; /* f32 (nsubx,nsubx) */
  int opticalBinning;
  dasp_cent_array centWeight;  /* f32 (corrsize,corrsize) or (nsubx,nsubx,corrsize,corrsize) or NULL */
  int correlationCentroiding;
  float corrThresh;
  dasp_cent_array corrPattern; /* f32 (nsubx,nsubx,C,C) in FFTW half-complex form (util/correlation.py:20-51) */
  dasp_cent_array corrimg;     /* f32 (nsubx,nsubx,C,C) output */
  int threshType, imageOnly;
  int useBrightest;            /* used when useBrightestArr.data == NULL */
  dasp_cent_array useBrightestArr; /* i32 (nsubaps,) */
  int preBinningFactor, parabolicFit, gaussianFit;
  float gaussianMinVal, gaussianReplaceVal;
  dasp_cent_array inputImage;  /* f32 2-D (nsubx*nimg)^2 or 4-D (nsubx,nsubx,nimg,nimg), or NULL */
  dasp_cent_array subapLocation; /* i32 nsubaps*4, or NULL */
} dasp_cent_params;

/* Message of the last failure on this thread ("" if none). */
const char *dasp_cent_last_error(void);
/* DASP_CENT_ABI_VERSION of the loaded library. */
int dasp_cent_abi_version(void);

/* cent.initialise (cmod/src/centmodule.c:2858-3161). Validates like the reference, uploads the
 * per-subaperture tables to the current HIP device and returns an opaque handle. */
int dasp_cent_initialise(const dasp_cent_params *p, dasp_cent_t **out);

/* cent.run (cmod/src/centmodule.c:3188-3229): phase -> images -> centroids for one simulation
 * step, including the nintegrations/nlatency cycle (:1774-1788, :3221-3223). Reads the borrowed
 * phs, writes cents/bimg/corrimg in place; *seconds receives the wall time like py_run. */
int dasp_cent_run(dasp_cent_t *h, float *seconds);

/* cent.runSlope (cmod/src/centmodule.c:3230-3272): detector image -> centroids. */
int dasp_cent_run_slope(dasp_cent_t *h, float *seconds);

/* cent.update (cmod/src/centmodule.c:2573-2856), split by the Python type of the new value.
 * _int: CALSOURCE, ADDPOISSON, NCEN, NTHREADS, OPTICALBINNING, CORRELATIONCENTROIDING,
 *       USEBRIGHTEST (scalar), INTEGSTEPS.
 * _float: SIG (scalar), READNOISE, READBG, NOISEFLOOR, SKYBRIGHTNESS (scalar), PXLPOWER, CORRTHRESH.
 * _array: SIG (array), SKYBRIGHTNESS (array), SPOTPSF, CENTWEIGHT, CORRPATTERN, REFCENTS, CALCOEFF,
 *         USEBRIGHTEST (array), OUTPUTPHASEARR; a NULL data pointer stands for None.
 * _caldata: CALDATA with (calData f32 (2,nsubx,nsubx,steps), calBounds i32 (2,nsubx,nsubx,2),
 *           calSteps f32 (steps,)); all NULL stands for None (:2153-2203). */
int dasp_cent_update_int(dasp_cent_t *h, int code, long value);
int dasp_cent_update_float(dasp_cent_t *h, int code, double value);
int dasp_cent_update_array(dasp_cent_t *h, int code, const dasp_cent_array *value);
int dasp_cent_update_caldata(dasp_cent_t *h, const dasp_cent_array *calData,
                             const dasp_cent_array *calBounds, const dasp_cent_array *calSteps);

/* cent.free (cmod/src/centmodule.c:3163-3186). */
int dasp_cent_free(dasp_cent_t *h);

/* cent.testcorrelation (cmod/src/centmodule.c:3275-3309): corr (n,n), a half-complex transformed
 * reference, is replaced by the correlation of img (n,n) with it. Host arrays. */
int dasp_cent_testcorrelation(const float *img, float *corr, int n);

/* ---- extensions (no counterpart in the reference) ---------------------------------------- */

/* Select the compute form (DASP_CENT_FORM_*). The DFT-GEMM forms cover nimg = 16 with n = 16 /
 * fftsize = clipsize = 32 and n = 17..32 / fftsize = clipsize = 64, the fused kernels' options
 * without brightest-pixel selection, and no spot kernel; other set-ups return
 * DASP_CENT_EUNSUPPORTED (a count set later runs the handle on the fused FP32 kernel until it is
 * cleared). The single-product form
 * (DASP_CENT_FORM_DFT_FP16_1) rounds each operand once to FP16 (no hi/lo compensation), also with
 * the detector options outside the inline detector (noise, sky array, relative floor, power,
 * weights). */
int dasp_cent_set_form(dasp_cent_t *h, int form);

/* Relative detector floor (no counterpart in the reference, whose readCCD knows threshType 0
 * and 1 only and leaves the data untouched for any other type, cmod/src/centmodule.c:785-817):
 * with enable != 0 and threshType 0, the floor subtracted and clipped at is noiseFloor * max(image)
 * when 0 < noiseFloor <= 1 (thresholdCorrelation's rule, :1488-1505), noiseFloor itself otherwise.
 * DASP_CENT_EINVAL when threshType is not 0. */
int dasp_cent_set_relative_floor(dasp_cent_t *h, int enable);

/* Phase factor exp(i(phi - t)) in the image kernels (fused, DFT-GEMM and the general path's image
 * formation): 1 (default) = the argument reduced to revolutions, r = x*c_hi - k + x*c_lo with
 * c_hi + c_lo = 1/(2 pi) (reduction error at most 2.6e-7 rad for |x| <= 400 rad), then the hardware
 * sine/cosine (v_sin_f32/v_cos_f32); 0 = sincosf (max abs error 7e-8). */
int dasp_cent_set_fast_trig(dasp_cent_t *h, int enable);

/* Launch grids of the single-WFS kernels: image_grid for the image kernel (fused or DFT-GEMM
 * form), corr_grid for the correlation/fit stage; 0 = the kernel's default persistent grid,
 * g > 0 = at most g workgroups, g < 0 = one workgroup per work unit (no cap), rounded up to a
 * multiple of 8 (one share per XCD). Every item is computed the same way whatever the grid: the
 * results do not depend on it. */
int dasp_cent_set_grid(dasp_cent_t *h, long image_grid, long corr_grid);
int dasp_cent_get_grid(dasp_cent_t *h, long *image_grid, long *corr_grid);

/* Autotuner of the single-WFS launches: times dasp_cent_run_device on the given device inputs
 * (one untimed run, then reps timed runs, HIP events on stream) for the image grids {default,
 * 256, 512, 1024, 2048, 4096 workgroups, uncapped}, keeps the fastest, then does the same for the
 * correlation/fit stage's grid when that stage runs; the choice stays in the handle
 * (dasp_cent_get_grid). Writes the slopes as dasp_cent_run_device does and synchronises stream.
 * *best_ms (may be NULL): the tuned step time in milliseconds. */
int dasp_cent_tune_grid(dasp_cent_t *h, const float *d_phs, long phs_stride, int nframes, float *d_cents,
                        const float *d_zpois, const float *d_zread, int reps, void *stream, double *best_ms);

/* Device-resident step over nframes phase frames (science/wfscent.py:630-702 keeps phase,
 * images and slopes in host memory; here they stay in HBM). Frame f reads
 * d_phs + f*phs_stride and writes d_cents + f*nsubaps*2, d_bimg + f*nsubaps*nimg^2 (if non-NULL)
 * and d_corrimg + f*nsubaps*C^2 (if non-NULL). d_zpois/d_zread are pre-drawn N(0,1) arrays of
 * nsubaps*nimg^2 per frame, 16-byte aligned (may be NULL when the noise they feed is off).
 * Asynchronous on stream.
 * Every frame is a readout: with nintegrations+nlatency > 1 the caller runs the integration
 * cycle (only the readout step forms an image, from that step's phase, as computeHll zeroes
 * its buffer on every call, cmod/src/centmodule.c:478). */
int dasp_cent_run_device(dasp_cent_t *h, const float *d_phs, long phs_stride, int nframes,
                         float *d_cents, float *d_bimg, float *d_corrimg,
                         const float *d_zpois, const float *d_zread, void *stream);

/* dasp_cent_run_device without any device allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments, kernels and results as dasp_cent_run_device;
 * the scratch that call would allocate in the handle (images handed to the correlation stage or
 * the general path when d_bimg is NULL, correlation images when d_corrimg is NULL, the resampled
 * phase grid, the general path's image-formation buffer) comes from one fixed device array of
 * DASP_CENT_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends. A call needing more
 * returns DASP_CENT_EINVAL (pass image buffers or fewer frames). Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (other streams, other threads, other handles) race. */
#define DASP_CENT_NOALLOC_SCRATCH_BYTES (64L << 20)
int dasp_cent_run_device_noalloc(dasp_cent_t *h, const float *d_phs, long phs_stride, int nframes,
                                 float *d_cents, float *d_bimg, float *d_corrimg,
                                 const float *d_zpois, const float *d_zread, void *stream);

/* Host noise generator used by dasp_cent_run: n standard normals from (seed, stream), PCG32
 * (O'Neill 2014) uniforms through the Box-Muller transform. Deterministic and host only. */
int dasp_cent_draw_normals(unsigned long seed, unsigned long stream, long n, float *out);

/* The same stream drawn on the device into d_out (n floats), asynchronously on stream (no
 * counterpart in the reference, which draws its noise inside the module). Every value equals the
 * host generator's except, rarely, in the last float bit (device vs host double-precision
 * log/sqrt/cos/sin). dasp_cent_run draws its readout noise this way. */
int dasp_cent_draw_normals_device(unsigned long seed, unsigned long stream, long n, float *d_out, void *stream_);

/* Complex spectrum (interleaved re,im; C*C per pattern) equivalent to a half-complex correlation
 * pattern, i.e. the K with calcCorrelation(img) = IDFT(DFT(pad(img)) * K) (host only). */
int dasp_cent_hc_to_spectrum(const float *pattern_hc, int C, float *spectrum);

/* Number of active subapertures, phase grid width and correlation image size of a handle. */
int dasp_cent_info(dasp_cent_t *h, int *nactive, int *npup, int *corrsize);

/* Kernel timing: when enabled, every later launch of the handle's main image kernel is bracketed
 * by HIP events on the launch stream. dasp_cent_kernel_time returns the summed milliseconds and
 * the number of launches measured since the last reset (it synchronises the last event). */
int dasp_cent_set_timing(dasp_cent_t *h, int enable);
int dasp_cent_kernel_time(dasp_cent_t *h, int reset, double *ms, long *launches);

/* ---- multi-WFS step (no counterpart in the reference: science/wfscent.py runs each WFS's
 * centcmod in turn, one host-threaded call per WFS and frame) -------------------------------- */

typedef struct dasp_cent_batch dasp_cent_batch_t;

/* One WFS of a multi-WFS step: a handle and its device buffers, laid out as for
 * dasp_cent_run_device (cents (nframes, nsubaps, 2); bimg, corrimg, zpois, zread optional). */
typedef struct {
  dasp_cent_t *handle;
  const float *phs;
  long phs_stride;
  int nframes;
  float *cents;
  float *bimg;
  float *corrimg;
  const float *zpois;
  const float *zread;
} dasp_cent_batch_item;

/* Build a step over nitems WFS (each handle at most once). WFS whose options the fused kernels
 * cover, in the FFT32 form, run in one persistent work-queue launch per (fftsize, phase pad, nimg,
 * trig mode); the others run their own kernels inside the same step graph. The buffers are
 * captured: a step always reads and writes the same device addresses. A handle changed by
 * update/set_form/set_fast_trig after this call is picked up at the next dasp_cent_batch_run
 * (the step graph is rebuilt). The handles must outlive the batch. */
int dasp_cent_batch_create(const dasp_cent_batch_item *items, int nitems, dasp_cent_batch_t **out);

/* Autotuner: for every persistent launch, time each (waves per workgroup in {1,2,4}) x
 * (static schedule with workgroups = k x CUs, k up to the resident limit, or one group per wave;
 * dynamic schedule claiming 2 or 8 groups per atomic at the resident limit) with HIP events over
 * reps launches on stream, keep the fastest, then rebuild the step graph. Synchronous. */
int dasp_cent_batch_tune(dasp_cent_batch_t *b, void *stream, int reps);

/* Run one step (a single hipGraphLaunch on stream; asynchronous). */
int dasp_cent_batch_run(dasp_cent_batch_t *b, void *stream);

/* Number of persistent launches in the step, and the configuration of launch i: its shape
 * (phase pad NP, fftsize, nimg, 1 if the WFS convolve with a spot kernel, 1 for fast trig), number of WFS, queue length in groups of 64/NP work items,
 * waves per workgroup, workgroups, schedule (0 static, c > 0 dynamic claims of c groups) and
 * the tuned time in ms (-1 before tuning). */
int dasp_cent_batch_launches(dasp_cent_batch_t *b);
int dasp_cent_batch_launch_info(dasp_cent_batch_t *b, int i, int shape[5], int *nwfs, long *groups,
                                int *waves_per_wg, long *workgroups, int *chunk, float *ms);

/* Set the configuration of launch i (e.g. one the autotuner chose in an earlier process) and
 * rebuild the step graph. */
int dasp_cent_batch_set_config(dasp_cent_batch_t *b, int i, int waves_per_wg, long workgroups, int chunk);

/* Number of WFS that run outside the persistent launches (their own kernels in the graph). */
int dasp_cent_batch_others(dasp_cent_batch_t *b);

int dasp_cent_batch_free(dasp_cent_batch_t *b);

/* ---- reconstructor step on the device (science/tomoRecon.py, dense reconmx) ------------------
 * The slopes never leave HBM: the slope vector of the reconstructor (length ncents: the x slopes of
 * every active sub-aperture of every WFS, then the y slopes; centIndex, science/tomoRecon.py:131-150,
 * getInput :682-700) is gathered on the device from a slope buffer (index[c] = position of entry c
 * within one frame of it, e.g. WFS k's slopes at a per-WFS offset), and each frame's actuator vector
 * is updated as
 *     acts = decay * acts - gains * (data . reconmx)        (:1140-1142, :1174)
 * layout 0: reconmx (ncents, nmodes) with row stride ld >= nacts, as science/tomoRecon.py's
 * dot(data, reconmx); layout 1: (nmodes, ncents) with ld >= ncents, its dot(reconmx, data). Only the
 * first nacts modes are formed. The matrix is device memory borrowed from the caller. */
typedef struct dasp_recon dasp_recon_t;
const char *dasp_recon_last_error(void);
int dasp_recon_create(int ncents, int nacts, int layout, const float *d_reconmx, long ld, const int *h_index,
                      dasp_recon_t **out);
/* Per-actuator gains and decay (host arrays of nacts, copied) or, for a NULL array, one value for
 * every actuator. Defaults: gain 1, decay 1. */
int dasp_recon_set_gains(dasp_recon_t *r, const float *h_gains, float gain, const float *h_decay, float decay);
/* nframes independent slope vectors: frame f reads d_slopes + f*slope_stride (slope_stride, the
 * floats per frame of the slope buffer, must exceed every index) and updates d_acts + f*acts_stride
 * (acts_stride >= nacts when nframes > 1). The matrix must hold (layout 0) ncents*ld or (layout 1)
 * nacts*ld floats. Asynchronous on stream; deterministic (fixed summation order). */
int dasp_recon_run(dasp_recon_t *r, const float *d_slopes, long slope_stride, int nframes, float *d_acts,
                   long acts_stride, void *stream);
int dasp_recon_free(dasp_recon_t *r);

#ifdef __cplusplus
}
#endif
#endif /* DASP_CENT_H */
