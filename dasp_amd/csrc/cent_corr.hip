// cent_corr.hip - centroid stage after the fused kernels for correlation centroiding and fits:
// calcCorrelation (FFT, mode 1) + thresholdCorrelation, then computeFit (parabolic, Gaussian
// or 3x3 quadratic interpolation) or computeCoG centred on the brightest pixel, then refCents
// (cmod/src/centmodule.c:1403-1505, :1069-1188, :873-971, :1899-1961).
//
// The fused image kernel writes each item's final detector image (after noise, threshold and
// pixel power); this kernel reads it back. One group of N lanes per item (N = corrsize, or nimg
// when there is no correlation), 64/N items per wave: lane r holds row r of the centre-padded
// image, transforms it in registers, the wave's LDS carries the two transposes, lane r then owns
// spectrum column r for the product with the pattern spectrum K (calcCorrelation(img) =
// IDFT2(DFT2(pad(img)) K), unnormalised, K built on the host from the half-complex pattern).
// The fit's window sums run in one lane in the order of the general kernel.
#include "cent_internal.h"
#include "fft_reg.h"
#include "wave_ops.h"

namespace dasp {

constexpr int kXcds = 8;

namespace {

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// (value, index) maximum with the first index winning ties; value 0 / index 0 when nothing is
// positive (the reference's scan starts from mx = 0 with a strict comparison).
template <int W>
__device__ __forceinline__ void gargmax(float &v, int &i) {
  auto step = [&](float ov, int oi) {
    if (ov > v || (ov == v && oi < i)) {
      v = ov;
      i = oi;
    }
  };
  auto movi = [](int x, auto ctrl) { return __builtin_amdgcn_mov_dpp(x, decltype(ctrl)::value, 0xF, 0xF, false); };
  if constexpr (W >= 2) step(dpp_mov<dpp::kQuadXor1>(v), movi(i, std::integral_constant<int, dpp::kQuadXor1>{}));
  if constexpr (W >= 4) step(dpp_mov<dpp::kQuadXor2>(v), movi(i, std::integral_constant<int, dpp::kQuadXor2>{}));
  if constexpr (W >= 8) step(dpp_mov<dpp::kRowHalfMirror>(v), movi(i, std::integral_constant<int, dpp::kRowHalfMirror>{}));
  if constexpr (W >= 16) step(dpp_mov<dpp::kRowMirror>(v), movi(i, std::integral_constant<int, dpp::kRowMirror>{}));
  if constexpr (W >= 32) step(__shfl_xor(v, 16), __shfl_xor(i, 16));
}

}  // namespace

template <int N, bool CORR>
__global__ __launch_bounds__(256) void corr_fit_kernel(CorrArgs a) {
  constexpr int IPW = 64 / N, TS = N + 1, WAVES = 4;
  __shared__ cf stage[WAVES][IPW * N * TS];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int sl = lane / N;
  const int r = lane - sl * N;
  cf *T = &stage[w][sl * N * TS];
  float *Tf = reinterpret_cast<float *>(T);
  // XCD-aware order: workgroup b runs on XCD b % 8, and XCD x takes the x-th eighth of the active
  // sub-apertures for every frame, so that its L2 keeps that slice's pattern spectra (read once
  // per item otherwise from HBM). Groups of IPW neighbouring sub-apertures of the slice, frame-major.
  const int xcd = blockIdx.x % kXcds;
  const int s0 = (int)((long)a.nactive * xcd / kXcds);
  const int cnt = (int)((long)a.nactive * (xcd + 1) / kXcds) - s0;
  const int gpf = (cnt + IPW - 1) / IPW;                               // groups per frame
  const long ngroups = cnt > 0 ? (a.nitems / a.nactive) * gpf : 0;    // this XCD's groups
  const long gstride = (long)(gridDim.x / kXcds) * WAVES;
  const int M = a.nimg;
  const int o = (N - M) / 2;   // centre padding when correlating (corrsize > nimg)
  struct Item {
    int frame, s;
    bool valid;
  };
  auto item_of = [&](long g) {
    Item it{0, 0, false};
    if (g < ngroups) {
      const long gf = g / gpf;
      const int rr = (int)(g - gf * gpf) * IPW + sl;
      if (rr < cnt) it = Item{(int)gf, a.active[s0 + rr], true};
    }
    return it;
  };
  // the lane's image row r (padded), read from a clamped in-bounds address and zeroed by a
  // select, so the next group's loads are issued unconditionally while this group is transformed
  auto load_row = [&](const Item &it, float (&v)[N]) {
    const float *img = a.img + (long)it.frame * a.img_stride + (long)it.s * M * M;
    const int ir = CORR ? r - o : r;
    const bool rin = it.valid && ir >= 0 && ir < M;
    const float *rp = img + (rin ? ir : 0) * M;
    static_for<0, N>([&](auto cc) {
      constexpr int c = decltype(cc)::value;
      const int ic = CORR ? c - o : c;
      const bool in = rin && ic >= 0 && ic < M;
      const float val = rp[in ? ic : 0];
      v[c] = in ? val : 0.f;
    });
  };
  long g = (long)(blockIdx.x / kXcds) * WAVES + w;
  Item cur = item_of(g);
  float crow[N];
  load_row(cur, crow);
  for (; g < ngroups; g += gstride) {
    const Item nxt = item_of(g + gstride);
    float nrow[N];
    load_row(nxt, nrow);
    const bool valid = cur.valid;
    const int frame = cur.frame, s = cur.s;
    const float *img = a.img + (long)frame * a.img_stride + (long)s * M * M;
    float row[N];
    if constexpr (CORR) {
      cf x[N];
      static_for<0, N>([&](auto cc) { x[decltype(cc)::value] = {crow[decltype(cc)::value], 0.f}; });
      fft_fwd<N>(x);
      static_for<0, N>([&](auto v) { T[r * TS + decltype(v)::value] = x[decltype(v)::value]; });
      wave_sync();
      cf y[N];
      static_for<0, N>([&](auto u) { y[decltype(u)::value] = T[decltype(u)::value * TS + r]; });
      wave_sync();
      fft_fwd<N>(y);
      const float2 *kp = a.corrK + (long)s * N * N + r;
      static_for<0, N>([&](auto u) {
        const float2 k = kp[decltype(u)::value * N];
        y[decltype(u)::value] = cmul(y[decltype(u)::value], {k.x, k.y});
      });
      fft_inv<N>(y);
      static_for<0, N>([&](auto u) { T[decltype(u)::value * TS + r] = y[decltype(u)::value]; });
      wave_sync();
      static_for<0, N>([&](auto v) { x[decltype(v)::value] = T[r * TS + decltype(v)::value]; });
      wave_sync();
      fft_inv<N>(x);
      static_for<0, N>([&](auto c) { row[decltype(c)::value] = x[decltype(c)::value].x; });
      // thresholdCorrelation (:1488-1505)
      float t = a.corrThresh;
      if (t <= 1.f && t > 0.f) {
        float mx = 0.f;
        static_for<0, N>([&](auto c) { mx = fmaxf(mx, row[decltype(c)::value]); });
        t = group_max<N>(mx) * t;
      }
      static_for<0, N>([&](auto c) {
        const float v = row[decltype(c)::value];
        row[decltype(c)::value] = v < t ? 0.f : v - t;
      });
    } else {
      static_for<0, N>([&](auto c) { row[decltype(c)::value] = crow[decltype(c)::value]; });
    }

    // sum and first maximum over the whole image (computeFit :1087-1120; computeCoG :941-969)
    float sum = 0.f, mv = 0.f;
    int mi = 0x7fffffff;
    static_for<0, N>([&](auto cc) {
      constexpr int c = decltype(cc)::value;
      sum += row[c];
      if (row[c] > mv) {
        mv = row[c];
        mi = r * N + c;
      }
    });
    sum = group_sum<N>(sum);
    gargmax<N>(mv, mi);
    const int mxpos = mv > 0.f ? mi : 0;
    float cx = 0.f, cy = 0.f;
    const int ncen = a.ncen;
    if (a.doFit && a.gaussianFit) {
      // computeFit takes the logarithm in place (:1087-1098) on the buffer it works on: the
      // correlation image when correlating, otherwise the detector image; both are returned
      // transformed
      static_for<0, N>([&](auto c) {
        const float v = row[decltype(c)::value];
        row[decltype(c)::value] = v < a.gaussianMinVal ? a.gaussianReplaceVal : logf(v);
      });
      if (!CORR && valid) {
        float *io = const_cast<float *>(img) + r * N;
        static_for<0, N>([&](auto c) { io[decltype(c)::value] = row[decltype(c)::value]; });
      }
    }
    if (CORR && a.corrimg && valid) {
      float *co = a.corrimg + (long)frame * a.corr_stride + (long)s * N * N + r * N;
      static_for<0, N>([&](auto c) { co[decltype(c)::value] = row[decltype(c)::value]; });
    }
    if (a.doFit) {
      static_for<0, N>([&](auto c) { Tf[r * N + decltype(c)::value] = row[decltype(c)::value]; });
      wave_sync();
      if (r == 0 && sum >= 0.f) {
        const int nc = (N - ncen) / 2;
        int fitsize = a.fitsize;
        const bool quad = fitsize < 0;
        if (quad || fitsize < 3) fitsize = 3;
        int startx = mxpos % N - fitsize / 2;
        int starty = mxpos / N - fitsize / 2;
        if (startx < nc) startx = nc;
        if (starty < nc) starty = nc;
        if (startx > nc + ncen - fitsize) startx = nc + ncen - fitsize;
        if (starty > nc + ncen - fitsize) starty = nc + ncen - fitsize;
        float x, y;
        if (!quad) {
          float vec[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
          for (int yy = 0; yy < fitsize; ++yy)
#pragma unroll 1
            for (int xx = 0; xx < fitsize; ++xx) {
              const float val = Tf[(yy + starty) * N + xx + startx];
              vec[0] += val * xx * xx;
              vec[1] += val * xx * yy;
              vec[2] += val * yy * yy;
              vec[3] += val * xx;
              vec[4] += val * yy;
              vec[5] += val;
            }
          float res[5];
          for (int i = 0; i < 5; ++i) {
            float t = 0.f;
            for (int j = 0; j < 6; ++j) t += a.fitMx[i * 6 + j] * vec[j];
            res[i] = t;
          }
          if (a.gaussianFit) {
            x = -res[3] / (2 * res[0]);
            y = -res[4] / (2 * res[2]);
          } else {
            x = (res[1] * res[4] / (2 * res[2]) - res[3]) / (2.f * res[0] - res[1] * res[1] / (2.f * res[2]));
            y = -(res[4] + res[1] * x) / (2.f * res[2]);
          }
        } else {
          const float *q = &Tf[starty * N + startx];
          const float a2 = (q[2 * N + 1] - q[1]) / 2.f;
          const float a3 = (q[2 * N + 1] - 2 * q[N + 1] + q[1]) / 2.f;
          const float a4 = (q[N + 2] - q[N]) / 2.f;
          const float a5 = (q[N + 2] - 2 * q[N + 1] + q[N]) / 2.f;
          const float a6 = (q[2 * N + 2] - q[2] - q[2 * N] + q[0]) / 4.f;
          y = (2 * a2 * a5 - a4 * a6) / (a6 * a6 - 4 * a3 * a5) + 1;
          x = (2 * a3 * a4 - a2 * a6) / (a6 * a6 - 4 * a3 * a5) + 1;
        }
        if (x < 0 || y < 0 || x >= fitsize || y >= fitsize) {
          x = (float)(mxpos % N - startx);
          y = (float)(mxpos / N - starty);
        }
        y -= ncen / 2.f - 0.5f;
        x -= ncen / 2.f - 0.5f;
        x += startx + ncen / 2.f - N / 2.f;
        y += starty + ncen / 2.f - N / 2.f;
        cx = x;
        cy = y;
      }
      wave_sync();
    } else {
      // computeCoG centred on the brightest pixel (corr mode 1), optionally weighted
      int sx = (int)truncf(mxpos % N - ncen / 2.f + 0.5f);
      int sy = (int)truncf(mxpos / N - ncen / 2.f + 0.5f);
      sx = min(max(sx, 0), N - ncen);
      sy = min(max(sy, 0), N - ncen);
      const float *wt = a.centWeight ? a.centWeight + (a.centWeightDim == 4 ? (long)s * N * N : 0L) : nullptr;
      float sw = 0.f, sxw = 0.f, syw = 0.f;
      if (r >= sy && r < sy + ncen) {
        static_for<0, N>([&](auto cc) {
          constexpr int c = decltype(cc)::value;
          if (c >= sx && c < sx + ncen) {
            const float v = wt ? row[c] * wt[r * N + c] : row[c];
            sw += v;
            sxw += v * c;
            syw += v * r;
          }
        });
      }
      sw = group_sum<N>(sw);
      sxw = group_sum<N>(sxw);
      syw = group_sum<N>(syw);
      if (sw > 0.f) {
        cx = sxw / sw + sx - (N - 1) / 2.f;
        cy = syw / sw + sy - (N - 1) / 2.f;
      }
    }
    if (valid && r == 0) {
      if (a.refCents) {
        cx -= a.refCents[s * 2];
        cy -= a.refCents[s * 2 + 1];
      }
      float *co = a.cents + (long)frame * a.cents_stride + s * 2;
      co[0] = cx;
      co[1] = cy;
    }
    cur = nxt;
    static_for<0, N>([&](auto c) { crow[decltype(c)::value] = nrow[decltype(c)::value]; });
  }
}

bool corr_fit_supported(int n) { return n == 8 || n == 16 || n == 32; }

hipError_t launch_corr_fit(const CorrArgs &a, int n, bool corr, hipStream_t st) {
  const long ipb = 4L * (64 / n);
  long blocks = (a.nitems + ipb - 1) / ipb;
  if (blocks <= 0) return hipSuccess;
  if (blocks > 2048) blocks = 2048;
  blocks = (blocks + kXcds - 1) / kXcds * kXcds;   // every XCD gets the same number of workgroups
#define DASP_CORR_CASE(NN)                                                                                  \
  if (n == NN) {                                                                                          \
    if (corr) hipLaunchKernelGGL((corr_fit_kernel<NN, true>), dim3((unsigned)blocks), dim3(256), 0, st, a);  \
    else hipLaunchKernelGGL((corr_fit_kernel<NN, false>), dim3((unsigned)blocks), dim3(256), 0, st, a);      \
    return hipGetLastError();                                                                            \
  }
  DASP_CORR_CASE(8)
  DASP_CORR_CASE(16)
  DASP_CORR_CASE(32)
#undef DASP_CORR_CASE
  return hipErrorInvalidValue;
}

}  // namespace dasp
