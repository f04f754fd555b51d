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

// computeFit (:1069-1188) on the staged N x N image Tf (row-major), by one lane: the window of
// fitsize around the first maximum mxpos, clamped to the ncen window, then the parabolic /
// Gaussian least-squares vertex or the 3x3 quadratic interpolation, the integer peak when the
// vertex leaves the window, and the reference's offsets. Window sums in the general kernel's order.
template <int N>
struct FitWindow {
  int fitsize, startx, starty;
  __device__ FitWindow(const CorrArgs &a, int mxpos) {
    const int ncen = a.ncen;
    const int nc = (N - ncen) / 2;
    fitsize = a.fitsize < 3 ? 3 : a.fitsize;   // 3 for the quadratic interpolation (fitsize -1)
    startx = mxpos % N - fitsize / 2;
    starty = mxpos / N - fitsize / 2;
    if (startx < nc) startx = nc;
    if (starty < nc) starty = nc;
    if (startx > nc + ncen - fitsize) startx = nc + ncen - fitsize;
    if (starty > nc + ncen - fitsize) starty = nc + ncen - fitsize;
  }
  // the vertex (x, y) in window coordinates to the reported offsets; the integer peak when the
  // vertex leaves the window
  __device__ void finish(const CorrArgs &a, int mxpos, float x, float y, float &cx, float &cy) const {
    const int ncen = a.ncen;
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
};

// the least-squares vertex from the six window sums (x^2, xy, y^2, x, y, 1 weighted by the values)
__device__ __forceinline__ void fit_vertex(const CorrArgs &a, const float (&vec)[6], float &x, float &y) {
  float res[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    float t = 0.f;
#pragma unroll
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
}

template <int N>
__device__ void fit_staged(const CorrArgs &a, const float *Tf, int mxpos, float &cx, float &cy) {
  const FitWindow<N> win(a, mxpos);
  const int fitsize = win.fitsize, startx = win.startx, starty = win.starty;
  const bool quad = a.fitsize < 0;
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
    fit_vertex(a, vec, x, y);
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
  win.finish(a, mxpos, x, y, cx, cy);
}

// computeFit's least-squares fit with the window sums spread over the N lanes of the group: lane r
// reads its row's fitsize values of the staged image and adds v, v xx, v xx^2; the six window sums are
// group totals of those, times yy and yy^2. The sums are taken in another order than the
// reference's serial loop, the same up to rounding. Result in every lane of the group.
template <int N>
__device__ void fit_group(const CorrArgs &a, const float *Tf, int r, int mxpos, float &cx, float &cy) {
  const FitWindow<N> win(a, mxpos);
  const int yy = r - win.starty;
  const float *q = Tf + r * N + win.startx;   // startx + fitsize <= N: the row stays in the image
  float p0 = 0.f, p1 = 0.f, p2 = 0.f, fk = 0.f;
  for (int k = 0; k < win.fitsize; ++k, fk += 1.f) {
    const float v = q[k];
    p0 += v;
    p1 += v * fk;
    p2 += v * (fk * fk);
  }
  const bool rin = yy >= 0 && yy < win.fitsize;
  const float fy = (float)yy;
  float vec[6];
  vec[0] = group_sum<N>(rin ? p2 : 0.f);
  vec[1] = group_sum<N>(rin ? p1 * fy : 0.f);
  vec[2] = group_sum<N>(rin ? p0 * (fy * fy) : 0.f);
  vec[3] = group_sum<N>(rin ? p1 : 0.f);
  vec[4] = group_sum<N>(rin ? p0 * fy : 0.f);
  vec[5] = group_sum<N>(rin ? p0 : 0.f);
  float x, y;
  fit_vertex(a, vec, x, y);
  win.finish(a, mxpos, x, y, cx, cy);
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
      if (a.fitsize > 0) {
        fit_group<N>(a, Tf, r, mxpos, cx, cy);
        if (!(sum >= 0.f)) cx = cy = 0.f;
      } else if (r == 0 && sum >= 0.f) {
        fit_staged<N>(a, Tf, mxpos, cx, cy);
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

// FFT correlation of two items per N-lane group (correlation mode 1). The two detector images A
// and B are real, so they share every transform: lane r transforms row r of A + iB; the row
// spectra are separated with the conjugate symmetry of a real row (A[v] = (Z[v] + conj Z[-v]) / 2,
// B[v] = (Z[v] - conj Z[-v]) / 2i, both held in the lane) and only the columns v = 0..N/2 of each
// are kept, the real columns 0 and N/2 of one image packed as one complex column: N columns for
// the pair, one per lane. Each lane transforms its column, multiplies by its image's pattern
// spectrum (for the packed column: separated, both halves multiplied, packed again: each half is
// conjugate-symmetric in u, so its inverse is real), transforms back; lane r then rebuilds row r of
// both inverse row spectra (the missing half by conjugate symmetry) and inverts them packed as
// one complex row: corr_A + i corr_B. Half the transforms of one item per group, the same
// arithmetic up to rounding. The threshold, sums and first maximum then run per image; the two
// fits run side by side in lanes 0 and 1 of the group on the staged correlation images.
template <int N, bool PAD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(N <= 8 ? 4 : 3))) void corr_pair_kernel(CorrArgs a) {
  constexpr int PPW = 64 / N, TS = N + 1, WAVES = 4, H = N / 2;
  __shared__ cf stage[WAVES][PPW * N * TS];
  static_assert(2 * N * N * sizeof(float) <= N * TS * sizeof(cf), "both fit images fit in the transform tile");
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int sl = lane / N;
  const int r = lane - sl * N;
  cf *T = &stage[w][sl * N * TS];
  // the fits' staged images reuse the tile once its last rows have been read back
  float *TfA = reinterpret_cast<float *>(T), *TfB = TfA + N * N;
  // XCD-aware order as in corr_fit_kernel; a pair is two neighbouring active sub-apertures of the
  // XCD's slice in one frame (the slice's last pair of a frame may hold one)
  const int xcd = blockIdx.x % kXcds;
  const int s0 = (int)((long)a.nactive * xcd / kXcds);
  const int cnt = (int)((long)a.nactive * (xcd + 1) / kXcds) - s0;
  const int ppf = (cnt + 1) / 2;                                         // pairs per frame
  const int gpf = (ppf + PPW - 1) / PPW;                                 // groups per frame
  const long ngroups = cnt > 0 ? (a.nitems / a.nactive) * gpf : 0;
  const long gstride = (long)(gridDim.x / kXcds) * WAVES;
  const int M = a.nimg;
  const int o = (N - M) / 2;
  struct Pair {
    int frame, sA, sB;
    bool vA, vB;
  };
  auto pair_of = [&](long g) {
    Pair p{0, 0, 0, false, false};
    if (g < ngroups) {
      const long gf = g / gpf;
      const int k = ((int)(g - gf * gpf) * PPW + sl) * 2;   // index of A in the slice
      if (k < cnt) {
        p.frame = (int)gf;
        p.sA = a.active[s0 + k];
        p.vA = true;
        if (k + 1 < cnt) {
          p.sB = a.active[s0 + k + 1];
          p.vB = true;
        } else {
          p.sB = p.sA;
        }
      }
    }
    return p;
  };
  auto load_row = [&](int frame, int s, bool valid, float (&v)[N]) {
    const float *img = a.img + (long)frame * a.img_stride + (long)s * M * M;
    if constexpr (!PAD) {   // nimg == N: no padding, one select per pixel (pair_of's idle slots
                            // point at sub-aperture 0 of frame 0, an in-bounds address)
      const float *rp = img + r * N;
      static_for<0, N>([&](auto cc) {
        const float val = rp[decltype(cc)::value];
        v[decltype(cc)::value] = valid ? val : 0.f;
      });
      return;
    }
    const int ir = r - o;
    const bool rin = valid && ir >= 0 && ir < M;
    const float *rp = img + (rin ? ir : 0) * M;
    static_for<0, N>([&](auto cc) {
      constexpr int c = decltype(cc)::value;
      const int ic = c - o;
      const bool in = rin && ic >= 0 && ic < M;
      const float val = rp[in ? ic : 0];
      v[c] = in ? val : 0.f;
    });
  };
  long g = (long)(blockIdx.x / kXcds) * WAVES + w;
  Pair cur = pair_of(g);
  float ra[N], rb[N];
  load_row(cur.frame, cur.sA, cur.vA, ra);
  load_row(cur.frame, cur.sB, cur.vB, rb);
  for (; g < ngroups; g += gstride) {
    const Pair nxt = pair_of(g + gstride);
    cf x[N];
    static_for<0, N>([&](auto cc) { x[decltype(cc)::value] = {ra[decltype(cc)::value], rb[decltype(cc)::value]}; });
    load_row(nxt.frame, nxt.sA, nxt.vA, ra);   // the next pair's rows, while this one is transformed
    load_row(nxt.frame, nxt.sB, nxt.vB, rb);
    const int frame = cur.frame;
    // ---- rows: one transform of A + iB, separated into the half spectra, staged as N slots:
    //      slot 0 = A[0] + i A[H], slots 1..H-1 = A[1..H-1], slot H = B[0] + i B[H],
    //      slots H+1..N-1 = B[1..H-1]
    fft_fwd<N>(x);
    static_for<0, H + 1>([&](auto vv) {
      constexpr int v = decltype(vv)::value, vm = (N - v) % N;
      const cf z = x[v], zm = x[vm];
      const cf av = {0.5f * (z.x + zm.x), 0.5f * (z.y - zm.y)};
      const cf bv = {0.5f * (z.y + zm.y), 0.5f * (zm.x - z.x)};
      if constexpr (v == 0) {
        T[r * TS + 0].x = av.x;
        T[r * TS + H].x = bv.x;
      } else if constexpr (v == H) {
        T[r * TS + 0].y = av.x;
        T[r * TS + H].y = bv.x;
      } else {
        T[r * TS + v] = av;
        T[r * TS + H + v] = bv;
      }
    });
    wave_sync();
    // ---- columns: lane r owns slot j = r
    const int j = r;
    const bool isB = j >= H;
    const int v = isB ? j - H : j;   // 0: the packed columns 0 and H
    cf y[N];
    static_for<0, N>([&](auto u) { y[decltype(u)::value] = T[decltype(u)::value * TS + j]; });
    wave_sync();
    fft_fwd<N>(y);
    // W[u] = alpha[u] Y[u] + beta[u] conj(Y[-u]) from the table's (alpha, beta): alpha = K[u][v],
    // beta = 0 on a regular column; on the packed column (v = 0) (K0 + KH) / 2 and (K0 - KH) / 2,
    // which multiply columns 0 and H by their own spectra and pack them again (one branch-free
    // form and one 16-byte load per value for all lanes)
    const float4 *kp = a.corrKp + (long)(isB ? cur.sB : cur.sA) * N * H + v;
    cf wv[N];
    static_for<0, N>([&](auto uu) {
      constexpr int u = decltype(uu)::value, um = (N - u) % N;
      const float4 ab = kp[u * H];
      const cf p = y[u], pm = y[um];
      wv[u] = {p.x * ab.x - p.y * ab.y + pm.x * ab.z + pm.y * ab.w, p.x * ab.y + p.y * ab.x + pm.x * ab.w - pm.y * ab.z};
    });
    static_for<0, N>([&](auto u) { y[decltype(u)::value] = wv[decltype(u)::value]; });
    fft_inv<N>(y);
    static_for<0, N>([&](auto u) { T[decltype(u)::value * TS + j] = y[decltype(u)::value]; });
    wave_sync();
    // ---- rows back: lane r rebuilds both inverse row spectra and inverts them packed
    cf q[N];
    {
      const cf p0 = T[r * TS + 0], pH = T[r * TS + H];
      q[0] = {p0.x, pH.x};   // A[0] + i B[0] (both real)
      q[H] = {p0.y, pH.y};   // A[H] + i B[H]
    }
    static_for<1, H>([&](auto vv) {
      constexpr int vq = decltype(vv)::value;
      const cf ga = T[r * TS + vq], gb = T[r * TS + H + vq];
      q[vq] = {ga.x - gb.y, ga.y + gb.x};         // GA + i GB
      q[N - vq] = {ga.x + gb.y, gb.x - ga.y};     // conj GA + i conj GB
    });
    wave_sync();
    fft_inv<N>(q);
    // ---- per image: threshold, sums, first maximum, output, fit or centred CoG
    float cxs[2] = {0.f, 0.f}, cys[2] = {0.f, 0.f};
    float sums[2];
    int mxs[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      float row[N];
      static_for<0, N>([&](auto c) { row[decltype(c)::value] = b ? q[decltype(c)::value].y : q[decltype(c)::value].x; });
      const bool valid = b ? cur.vB : cur.vA;
      const int s = b ? cur.sB : cur.sA;
      // thresholdCorrelation (:1488-1505)
      float t = a.corrThresh;
      if (t <= 1.f && t > 0.f) {
        float mx = 0.f;
        static_for<0, N>([&](auto c) { mx = fmaxf(mx, row[decltype(c)::value]); });
        t = group_max<N>(mx) * t;
      }
      static_for<0, N>([&](auto c) {
        const float vv = row[decltype(c)::value];
        row[decltype(c)::value] = vv < t ? 0.f : vv - t;
      });
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
      sums[b] = sum;
      mxs[b] = mxpos;
      if (a.doFit && a.gaussianFit) {
        // computeFit's logarithm in place (:1087-1098): the returned correlation image holds it
        static_for<0, N>([&](auto c) {
          const float vv = row[decltype(c)::value];
          row[decltype(c)::value] = vv < a.gaussianMinVal ? a.gaussianReplaceVal : logf(vv);
        });
      }
      if (a.corrimg && valid) {
        float *co = a.corrimg + (long)frame * a.corr_stride + (long)s * N * N + r * N;
        static_for<0, N>([&](auto c) { co[decltype(c)::value] = row[decltype(c)::value]; });
      }
      if (a.doFit) {
        float *Tf = b ? TfB : TfA;
        static_for<0, N>([&](auto c) { Tf[r * N + decltype(c)::value] = row[decltype(c)::value]; });
      } else {
        // computeCoG centred on the brightest pixel, optionally weighted
        const int ncen = a.ncen;
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
              const float vv = wt ? row[c] * wt[r * N + c] : row[c];
              sw += vv;
              sxw += vv * c;
              syw += vv * r;
            }
          });
        }
        sw = group_sum<N>(sw);
        sxw = group_sum<N>(sxw);
        syw = group_sum<N>(syw);
        if (sw > 0.f) {
          cxs[b] = sxw / sw + sx - (N - 1) / 2.f;
          cys[b] = syw / sw + sy - (N - 1) / 2.f;
        }
      }
    }
    if (a.doFit) {
      wave_sync();
      if (a.fitsize > 0) {
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          fit_group<N>(a, b ? TfB : TfA, r, mxs[b], cxs[b], cys[b]);
          if (!(sums[b] >= 0.f)) cxs[b] = cys[b] = 0.f;
        }
      } else if (r < 2 && (r ? sums[1] : sums[0]) >= 0.f) {
        // the quadratic interpolation: lane 0 of the group for A, lane 1 for B
        float fx, fy;
        fit_staged<N>(a, r ? TfB : TfA, r ? mxs[1] : mxs[0], fx, fy);
        cxs[0] = cxs[1] = fx;
        cys[0] = cys[1] = fy;
      }
      wave_sync();
    }
    // lane 0 of the group reports A, lane 1 reports B
    const int bsel = r;
    float cx = bsel ? cxs[1] : cxs[0], cy = bsel ? cys[1] : cys[0];
    const bool valid = bsel ? cur.vB : cur.vA;
    if (r < 2 && valid) {
      const int s = bsel ? cur.sB : cur.sA;
      if (a.refCents) {
        cx -= a.refCents[s * 2];
        cy -= a.refCents[s * 2 + 1];
      }
      float *co = a.cents + (long)frame * a.cents_stride + s * 2;
      co[0] = cx;
      co[1] = cy;
    }
    cur = nxt;
  }
}

bool corr_fit_supported(int n) { return n == 8 || n == 16 || n == 32; }

hipError_t launch_corr_fit(const CorrArgs &a, int n, bool corr, hipStream_t st, long grid) {
  const long ipb = 4L * (64 / n);
  if (a.nitems <= 0) return hipSuccess;
  const long blocks = launch_grid((a.nitems + ipb - 1) / ipb, 2048, grid);
  // correlation (mode 1): two items per lane group through shared transforms (corr_pair_kernel)
  if (corr && a.corrKp && (n == 8 || n == 16)) {
    const long ppb = 4L * (64 / n), pairs = (a.nitems + 1) / 2 + kXcds;
    const long pb = launch_grid((pairs + ppb - 1) / ppb, 2048, grid);
    const bool pad = a.nimg != n;
    if (n == 8) {
      if (pad) hipLaunchKernelGGL((corr_pair_kernel<8, true>), dim3((unsigned)pb), dim3(256), 0, st, a);
      else hipLaunchKernelGGL((corr_pair_kernel<8, false>), dim3((unsigned)pb), dim3(256), 0, st, a);
    } else {
      if (pad) hipLaunchKernelGGL((corr_pair_kernel<16, true>), dim3((unsigned)pb), dim3(256), 0, st, a);
      else hipLaunchKernelGGL((corr_pair_kernel<16, false>), dim3((unsigned)pb), dim3(256), 0, st, a);
    }
    return hipGetLastError();
  }
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
