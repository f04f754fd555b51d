// cent_form.hip - image formation on gfx950: phase -> pup*exp(i(phi-t)) -> zero-padded 2-D FFT
// -> |X|^2 placed as the reference's high-light-level image "hll".
//
// Follows computeHll + addPowerSpecCentral (cmod/src/centmodule.c:430-509). Layout: a wave owns
// SPW = 64/NP subapertures; lane (s, j) loads column j of subaperture s with coalesced reads of
// the phase rows, transforms it along i in registers (NP-point FFTs, the zero half of the F-point
// input is never touched), the wave transposes through LDS in R = F/NP phases (phase r stages the
// rows k = q*R + r, so that NP x NP values per subaperture sit in LDS at a time), and lane (s, m)
// transforms row k = m*R + r along j, again in registers.
#include "cent_internal.h"
#include "fft_reg.h"

namespace dasp {

template <int NP, int F>
struct FormShape {
  static constexpr int SPW = 64 / NP;          // subapertures per wave
  static constexpr int R = F / NP;             // rows per lane in the second pass
  static constexpr int TS = NP + 1;            // transpose row stride (complex), padded
  static constexpr int WAVE_LDS = SPW * NP * TS * (int)sizeof(cf);
  static constexpr int W = (65536 / WAVE_LDS) >= 4 ? 4 : ((65536 / WAVE_LDS) >= 2 ? 2 : 1);
};

// Column transform: lane holds E[0..NP-1][j] and gets Z[k][j], k < F.
template <int NP, int F, bool FAST>
__device__ __forceinline__ void load_column(const float *__restrict__ ph, const float *__restrict__ am,
                                            const float *__restrict__ pu, const float *__restrict__ tilt,
                                            int n, int npup, int j, bool valid, cf (&x)[NP]) {
  static_for<0, NP>([&](auto ii) {
    constexpr int i = decltype(ii)::value;
    cf v = {0.f, 0.f};
    if (valid && i < n && j < n) {
      const float phi = ph[(long)i * npup + j];
      const float th = phi - tilt[i * n + j];   // float subtraction, as the reference (:484)
      float sn, cs;
      phase_sincos<FAST>(th, &sn, &cs);
      float p = pu[i * n + j];
      if (am) p *= am[(long)i * npup + j];
      v = {p * cs, p * sn};
    }
    x[i] = v;
  });
}

template <int NP, int F, bool FAST>
__global__ __launch_bounds__((64 * FormShape<NP, F>::W)) void form_hll_kernel(FormArgs a) {
  using S = FormShape<NP, F>;
  constexpr int SPW = S::SPW, R = S::R, TS = S::TS, W = S::W;
  __shared__ cf tbuf[W][SPW * NP * TS];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int sl = lane / NP;
  const int j = lane - sl * NP;
  const long item = ((long)blockIdx.x * W + w) * SPW + sl;
  const bool valid = (sl < SPW) && item < a.nitems;
  int frame = 0, s = 0;
  if (valid) {
    frame = (int)(item / a.nactive);
    s = a.active[item - (long)frame * a.nactive];
  }
  const int n = a.n;
  const int sy = s / a.nsubx, sx = s - (s / a.nsubx) * a.nsubx;
  const long off = (long)frame * a.phs_stride + (long)(sy * n) * a.npup + sx * n;
  const float *ph = a.phs + off;
  const float *am = a.amp ? a.amp + off : nullptr;
  const float *pu = a.pup + (long)s * n * n;

  cf x[NP];
  load_column<NP, F, FAST>(ph, am, pu, a.tilt, n, a.npup, j, valid, x);
  cf z[F];
  fft_fwd_zp<NP, F>(x, z);
  cf *tb = &tbuf[w][sl * NP * TS];

  const int Q = a.padded;
  float *hll = a.hll + item * (long)Q * Q;
  static_for<0, R>([&](auto rr) {
    constexpr int r = decltype(rr)::value;
    // phase r: rows k = q*R + r of the column spectra through the wave's LDS
    static_for<0, NP>([&](auto qq) { tb[decltype(qq)::value * TS + j] = z[decltype(qq)::value * R + r]; });
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int k = j * R + r;
    cf row[NP];
    static_for<0, NP>([&](auto jj) { row[decltype(jj)::value] = tb[j * TS + decltype(jj)::value]; });
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    cf X[F];
    fft_fwd_zp<NP, F>(row, X);
    if (!valid) return;
    if (Q <= F) {
      const int o = (F - Q) >> 1;
      if (k >= o && k < o + Q) {
        float *dst = hll + (long)(Q - 1 - (k - o)) * Q;
        if ((Q & 3) == 0 && (o & 3) == 0) {
          // four rotated values per 16-byte store (the row's pixels l = o+4m .. o+4m+3 land at
          // Q-4(m+1) .. Q-1-4m), a quarter of the store instructions of one value each
          static_for<0, F / 4>([&](auto mm) {
            constexpr int m = decltype(mm)::value, l = 4 * m;
            if (l >= o && l < o + Q)
              *reinterpret_cast<float4 *>(dst + Q - 4 - (l - o)) =
                  make_float4(cabs2(X[l + 3]), cabs2(X[l + 2]), cabs2(X[l + 1]), cabs2(X[l]));
          });
        } else {
          static_for<0, F>([&](auto ll) {
            constexpr int l = decltype(ll)::value;
            if (l >= o && l < o + Q) dst[Q - 1 - (l - o)] = cabs2(X[l]);
          });
        }
      }
    } else {
      const int rs = (Q - F) >> 1;
      float *dst = hll + (long)(F - 1 - k + rs) * Q;
      static_for<0, F>([&](auto ll) {
        constexpr int l = decltype(ll)::value;
        dst[F - 1 - l + rs] = cabs2(X[l]);
      });
    }
  });
}

template <int NP, int F>
static hipError_t launch_form_t(const FormArgs &a, hipStream_t st) {
  using S = FormShape<NP, F>;
  const long per_block = (long)S::W * S::SPW;
  const long blocks = (a.nitems + per_block - 1) / per_block;
  if (blocks <= 0) return hipSuccess;
  if (a.fast_trig) hipLaunchKernelGGL((form_hll_kernel<NP, F, true>), dim3((unsigned)blocks), dim3(64 * S::W), 0, st, a);
  else hipLaunchKernelGGL((form_hll_kernel<NP, F, false>), dim3((unsigned)blocks), dim3(64 * S::W), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_form(const FormArgs &a, int np, int F, hipStream_t st) {
#define DASP_FORM_CASE(A, B) \
  if (np == A && F == B) return launch_form_t<A, B>(a, st);
  DASP_FORM_CASE(8, 8)
  DASP_FORM_CASE(8, 16)
  DASP_FORM_CASE(8, 32)
  DASP_FORM_CASE(16, 16)
  DASP_FORM_CASE(16, 32)
  DASP_FORM_CASE(16, 64)
  DASP_FORM_CASE(32, 32)
  DASP_FORM_CASE(32, 64)
#undef DASP_FORM_CASE
  return hipErrorInvalidValue;
}

__global__ void fill_kernel(float *p, long n, float v) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

hipError_t launch_fill(float *p, long n, float v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  long blocks = (n + 255) / 256;
  if (blocks > 65535) blocks = 65535;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

}  // namespace dasp
