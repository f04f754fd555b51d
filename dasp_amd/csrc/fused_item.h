// fused_item.h - device code of the fused FP32 path shared by the single-WFS kernel
// (cent_fused.hip) and the multi-WFS persistent kernel (cent_batch.hip): phase -> |FFT|^2 ->
// binned detector image -> noise -> threshold -> CoG for one work item per NP-lane group.
//
// Layout (wave64): a wave owns SPW = 64/NP subapertures, lane (s, j) owns phase column j of
// subaperture s. Pass 1 transforms columns along i in registers. The transpose to rows goes
// through the wave's private LDS region in R = F/NP phases of NP rows each (so only NP x NP
// complex values per subaperture are staged at a time), and pass 2 transforms row k = m*R + r
// along j. |X|^2 is binned straight from registers: with clip == F the 180-degree rotation maps
// row k to detector row (F-1-k)/B and column l to (F-1-l)/B, all compile-time indices.
#pragma once
#include "cent_internal.h"
#include "fft_reg.h"
#include "noise.h"
#include "wave_ops.h"

namespace dasp {

template <int NP, int F, int B>
struct FusedShape {
  static constexpr int SPW = 64 / NP;
  static constexpr int R = F / NP;
  static constexpr int NIMG = F / B;
  static constexpr int RB = R >= B ? R / B : 1;   // detector rows held per lane
  static constexpr int G = B >= R ? B / R : 1;    // lanes that share one detector row
  static constexpr int TS = NP + 1;               // staged row stride (complex)
  static constexpr int WAVES = 4;
  static constexpr int PX = RB * NIMG;            // detector pixels per lane
  static constexpr int PXS = PX + 1;              // their LDS stride per lane (odd: no bank conflicts)
  static_assert(R * NP == F && NIMG * B == F, "shape");
  static_assert((R % B == 0) || (B % R == 0), "binning must align with rows per lane");
};

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int NP>
struct PhaseCol {
  float ph[NP], pu[NP];
};

// A work item as a lane group sees it: frame, subaperture index, and whether the lane group
// holds a real item (padding groups compute on zeros and write nothing).
struct ItemRef {
  int frame;
  int s;
  int off;   // the phase window's offset in its frame, (sy*n)*npup + sx*n
  bool valid;
};

__device__ __forceinline__ long wave_uniform(unsigned long long v) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (long)(((unsigned long long)hi << 32) | lo);
}

// Item idx (index into the active list) of frame `frame`; idx >= nactive or frame < 0 is padding.
// Padding reads item 0 (unconditional loads, no branch) and is never written.
__device__ __forceinline__ ItemRef item_at(const FusedArgs &a, int frame, int idx) {
  ItemRef r;
  r.valid = frame >= 0 && idx < a.nactive;
  const int ci = r.valid ? idx : 0;
  r.frame = r.valid ? frame : 0;
  r.s = a.active[ci];
  r.off = a.winoff[ci];
  return r;
}

// Item first + sl of the frame-major numbering frame * nactive + idx, where `first` (the group's
// first item) is the same for the whole wave: the division runs once, on the scalar unit.
__device__ __forceinline__ ItemRef item_flat(const FusedArgs &a, long first, int sl) {
  first = wave_uniform((unsigned long long)first);
  const long f0 = first / a.nactive;
  int idx = (int)(first - f0 * a.nactive) + sl;
  int f = (int)f0;
  if (idx >= a.nactive) {
    idx -= a.nactive;
    ++f;
  }
  const bool in = first + sl < a.nitems;
  return item_at(a, in ? f : -1, idx);
}

// Item first + sl of the subaperture-major numbering idx * nframes + frame, for the multi-WFS
// spot kernels: the slices of a group are then consecutive frames of one subaperture and their
// loads of its spot-spectrum rows and pupil hit the same cache lines.
__device__ __forceinline__ ItemRef item_flat_by_subap(const FusedArgs &a, long first, int sl) {
  first = wave_uniform((unsigned long long)first);
  const long nfr = a.nitems / a.nactive;
  const long i0 = first / nfr;
  int f = (int)(first - i0 * nfr) + sl;
  int idx = (int)i0;
  while (f >= nfr) {   // at most sl / nfr + 1 steps
    f -= (int)nfr;
    ++idx;
  }
  const bool in = first + sl < a.nitems;
  return item_at(a, in ? f : -1, idx);
}

// Load phase column j (rows i < n) and the pupil of one work item. The loads are unconditional
// (no per-element branches): with n < NP the padding rows/columns read a clamped in-window
// address and the pupil factor is zeroed by a product, so E is exactly 0 there.
// With UF the caller's frame fw is the same for the whole wave (the fused kernel's order): the
// row bases are then scalar and each load takes one 32-bit lane offset (no 64-bit address
// arithmetic per element).
template <int NP, bool UF = false>
__device__ __forceinline__ void load_item(const FusedArgs &a, const ItemRef &it, int j, PhaseCol<NP> &c, int fw = 0) {
  const int n = a.n;
  const float *ph = a.phs + (long)it.frame * a.phs_stride + it.off;
  const float *pu = a.pup + (long)it.s * n * n;
  if (n == NP) {
    if constexpr (UF) {
      const float *base = a.phs + (long)fw * a.phs_stride;
      // 32-bit unsigned byte offsets (zero-extended), so that the loads take the scalar-base
      // plus vector-offset form
      const unsigned po = (unsigned)(it.off + j) * 4u, uo = (unsigned)(it.s * (NP * NP) + j) * 4u;
      static_for<0, NP>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        c.ph[i] = *(const float *)((const char *)(base + (long)i * a.npup) + po);
        c.pu[i] = *(const float *)((const char *)(a.pup + i * NP) + uo);
      });
    } else {
      static_for<0, NP>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        c.ph[i] = ph[(long)i * a.npup + j];
        c.pu[i] = pu[i * NP + j];
      });
    }
  } else if constexpr (UF && NP <= 16) {
    // scalar row bases (clamped rows ic are wave-uniform) and one lane offset per operand, as
    // above: nothing per row to hold across items but the uniform row index (for NP = 32 the
    // 32 hoisted row bases cost more in scalar registers than the per-lane form)
    const int jc = j < n ? j : n - 1;
    const float *base = a.phs + (long)fw * a.phs_stride;
    const unsigned po = (unsigned)(it.off + jc) * 4u, uo = (unsigned)(it.s * n * n + jc) * 4u;
    const float jm = j < n ? 1.f : 0.f;
    static_for<0, NP>([&](auto ii) {
      constexpr int i = decltype(ii)::value;
      const int ic = i < n ? i : n - 1;
      c.ph[i] = *(const float *)((const char *)(base + (long)ic * a.npup) + po);
      c.pu[i] = *(const float *)((const char *)(a.pup + ic * n) + uo) * (i < n ? jm : 0.f);
    });
  } else {
    const int jc = j < n ? j : n - 1;
    static_for<0, NP>([&](auto ii) {
      constexpr int i = decltype(ii)::value;
      const int ic = i < n ? i : n - 1;
      c.ph[i] = ph[(long)ic * a.npup + jc];
      c.pu[i] = pu[ic * n + jc] * ((i < n && j < n) ? 1.f : 0.f);
    });
  }
}

// LDS of one wave of the fused kernels, in complex values. During the transforms the region holds
// SPW subaperture slots of fft_slot_cf() values: the pass-1 transpose (NP rows of NP + 1), or with
// a spot kernel the F x (NP + 1) spectrum transpose plus one value so consecutive subapertures
// start on different banks. The general detector then reuses the region for the 64 lanes'
// pixels (PXS floats each; up to 64 pixels per lane when nimg == fftsize, more than one slot).
template <int NP, int F, bool SPOT>
constexpr int fft_slot_cf() {
  return SPOT ? F * (NP + 1) + 1 : NP * (NP + 1);
}
template <int NP, int F, int B>
constexpr int detector_cf() {
  return (64 * FusedShape<NP, F, B>::PXS + 1) / 2;
}
template <int NP, int F, int B, bool SPOT>
constexpr int wave_lds_cf() {
  constexpr int t = (64 / NP) * fft_slot_cf<NP, F, SPOT>(), d = detector_cf<NP, F, B>();
  return t > d ? t : d;
}

// LGS spot convolution, doConvolution (cmod/src/centmodule.c:524-564): the reference convolves
// the 180-degree rotated |X|^2 (hll) with the quadrant-swapped spot through r2c/c2r transforms
// (unnormalised, so the result carries psfsize^2, absorbed by the theoretical total). With
// R a -> a(F-1-.) in both axes, conv(R a, g) = R conv(a, g(-.)) and the spectrum of g(-.) is
// conj(OTF) for a real g: the unrotated |X|^2 of the lane's rows k0 = 2j, k1 = 2j+1 (z[l].x,
// z[l].y) is convolved with conj(OTF) and the rotated binning map is kept.
// The detector bins 2 x 2 (F = 2 nimg), and the block sum over a pixel pair (2m, 2m+1) of an
// unnormalised inverse transform is the H-point inverse transform (H = F/2) of the folded
// spectrum Y[u] f(u) + Y[u+H] f(u+H), f(u) = 1 + exp(2 pi i u / F). The host multiplies both fold
// factors into the kernel table, so the inverse transforms are H-point ones that yield the binned
// values directly:
//   rows:    one F-point FFT of the packed pair, Hermitian unpack in registers (v = 0..F/2);
//   columns: through LDS, lane m owns column v = m (m >= 1) or the packed pair v = 0 (re) and
//            v = F/2 (im) (m = 0), one F-point FFT, times the table, fold, one H-point inverse;
//   rows:    back through LDS, lane j rebuilds binned row j's folded spectrum from its Hermitian
//            half, one H-point inverse: the row's nimg binned values.
// Kernel table per subaperture (host-built from OTF, K = conj OTF): row m < NP holds
// A_m(u) = K(u,m) f(u) f(m) for m >= 1 and K(u,0) f(u) for m = 0 (row NP, the same as row 0, is
// not read here). Lane m = 0 forms A_0 (G(u) + conj G(-u)) = 2 K(u,0) f(u) S0(u) for the Hermitian
// column S0 packed in G (the column F/2 drops out: f(F/2) = 0).
template <int NP, int F, int B, int RB, int NIMG>
__device__ __forceinline__ void spot_convolve(const FusedArgs &a, int s, int j, cf (&z)[F], cf *T,
                                              float (&acc)[RB][NIMG]) {
  static_assert(F == 2 * NP, "one column per lane needs F == 2 NP");
  static_assert(B == 2 && RB == 1 && NIMG == F / 2, "the folded inverse transforms bin 2 x 2");
  constexpr int TS2 = NP + 1, H = F / 2;
  const int k0 = 2 * j;
  fft_fwd<F>(z);
  cf p0{0.f, 0.f}, p1{0.f, 0.f};
  static_for<0, H + 1>([&](auto vv) {
    constexpr int v = decltype(vv)::value;
    const cf zv = z[v], zc = z[(F - v) % F];
    const cf h0 = {0.5f * (zv.x + zc.x), 0.5f * (zv.y - zc.y)};    // row k0: (zv + conj zc) / 2
    const cf h1 = {0.5f * (zv.y + zc.y), -0.5f * (zv.x - zc.x)};   // row k1: (zv - conj zc) / 2i
    if constexpr (v == 0) {
      p0.x = h0.x;
      p1.x = h1.x;
    } else if constexpr (v == H) {
      p0.y = h0.x;
      p1.y = h1.x;
      T[k0 * TS2] = p0;
      T[(k0 + 1) * TS2] = p1;
    } else {
      T[k0 * TS2 + v] = h0;
      T[(k0 + 1) * TS2 + v] = h1;
    }
  });
  wave_lds_sync();
  cf c[F];
  static_for<0, F>([&](auto kk) { c[decltype(kk)::value] = T[decltype(kk)::value * TS2 + j]; });
  wave_lds_sync();
  const float2 *kp = a.otfk + (a.otf_per_subap ? (long)s : 0L) * ((NP + 1) * F);
  fft_fwd<F>(c);
  cf o[F];
  // lane 0's packed column: A_0 G(u) + A_NP conj(G(-u)) = A_0 (G(u) + conj G(-u)) (A_NP = A_0), so
  // every lane forms (G(u) + m conj G(-u)) times its table row, m = 1 on lane 0 and 0 elsewhere
  const float m0 = j == 0 ? 1.f : 0.f;
  static_for<0, H + 1>([&](auto uu) {
    constexpr int u = decltype(uu)::value, um = (F - u) % F;
    const cf gu = c[u], gm = c[um];
    const cf su = {fmaf(m0, gm.x, gu.x), fmaf(-m0, gm.y, gu.y)};
    const cf sm = {fmaf(m0, gu.x, gm.x), fmaf(-m0, gu.y, gm.y)};
    const float2 au = kp[j * F + u], am = kp[j * F + um];
    o[u] = cmul(su, {au.x, au.y});
    if constexpr (um != u) o[um] = cmul(sm, {am.x, am.y});
  });
  // fold in u and invert: binned row pairs m' = 0..H-1 of column j
  cf of[H];
  static_for<0, H>([&](auto uu) { of[decltype(uu)::value] = cadd(o[decltype(uu)::value], o[decltype(uu)::value + H]); });
  fft_inv<H>(of);
  static_for<0, H>([&](auto mm) { T[decltype(mm)::value * TS2 + j] = of[decltype(mm)::value]; });
  wave_lds_sync();
  // binned row j: G[v] for v = 0..H-1 from the columns (G[0] real, G[H] = 0 after the fold
  // factor), the other half by conjugate symmetry; folded: Gf[v] = G[v] + conj(G[H - v])
  cf q[H];
  q[0] = {T[j * TS2].x, 0.f};
  static_for<1, H>([&](auto vv) {
    constexpr int v = decltype(vv)::value;
    const cf g = T[j * TS2 + v], gc = T[j * TS2 + H - v];
    q[v] = {g.x + gc.x, g.y - gc.y};
  });
  wave_lds_sync();
  fft_inv<H>(q);
  static_for<0, H>([&](auto nn) {
    constexpr int n = decltype(nn)::value;
    acc[0][H - 1 - n] += q[n].x;
  });
}

// One work item (frame, subaperture) per NP-lane group: image formation, detector, CoG.
// wl: the wave's LDS region (wave_lds_cf values).
template <int NP, int F, int B, bool FAST, bool SPOT = false>
__device__ __forceinline__ void fused_item(const FusedArgs &a, const ItemRef &it, int j, int w, int sl,
                                           const PhaseCol<NP> &col, cf *wl) {
  using S = FusedShape<NP, F, B>;
  constexpr int R = S::R, NIMG = S::NIMG, RB = S::RB, G = S::G, TS = S::TS, PX = S::PX, PXS = S::PXS;
  cf *st = wl + sl * fft_slot_cf<NP, F, SPOT>();   // this lane group's subaperture slot
  const bool valid = it.valid;
  const int frame = it.frame, s = it.s;
  const int n = a.n;
  // ---- pass 1 operand: E[i][j] = pup * exp(i (phi - t)) ----
  cf x[NP];
  static_for<0, NP>([&](auto ii) {
    constexpr int i = decltype(ii)::value;
    const float th = col.ph[i] - a.tiltstep * (float)(i + j + 1 - n);   // float ops as the reference
    float sn, cs;
    phase_sincos<FAST>(th, &sn, &cs);
    x[i] = {col.pu[i] * cs, col.pu[i] * sn};
  });

  // ---- pass 1, transpose in R phases, pass 2, binning ----
  // Pass 1 is one zero-padded F-point transform of column j (Z[k][j], k < F). Phase r stages
  // rows k = q*R + r (q < NP), so that only NP x NP values per subaperture sit in LDS at a time;
  // lane j then transforms row k = j*R + r along its NP nonzero columns, zero-padded to F.
  float acc[RB][NIMG];
  static_for<0, RB>([&](auto ib) {
    static_for<0, NIMG>([&](auto jb) { acc[decltype(ib)::value][decltype(jb)::value] = 0.f; });
  });
  cf zs[SPOT ? F : 1];   // spot kernel: |X|^2 of rows 2j (re) and 2j+1 (im), unrotated
  cf z1[F];
  fft_fwd_zp<NP, F>(x, z1);
  static_for<0, R>([&](auto rr) {
    constexpr int r = decltype(rr)::value;
    static_for<0, NP>([&](auto qq) { st[decltype(qq)::value * TS + j] = z1[decltype(qq)::value * R + r]; });
    wave_lds_sync();
    cf row[NP];
    static_for<0, NP>([&](auto jj) { row[decltype(jj)::value] = st[j * TS + decltype(jj)::value]; });
    wave_lds_sync();
    constexpr int ib = R >= B ? r / B : 0;
    cf y2[F];
    fft_fwd_zp<NP, F>(row, y2);
    static_for<0, F>([&](auto ll) {
      constexpr int l = decltype(ll)::value;
      constexpr int jb = (F - 1 - l) / B;
      if constexpr (SPOT) {
        if constexpr (r == 0) zs[l].x = cabs2(y2[l]);
        else zs[l].y = cabs2(y2[l]);
      } else {
        acc[ib][jb] = fmaf(y2[l].y, y2[l].y, fmaf(y2[l].x, y2[l].x, acc[ib][jb]));   // |X|^2 binned: two fmas
      }
    });
  });
  if constexpr (SPOT) spot_convolve<NP, F, B, RB, NIMG>(a, s, j, zs, st, acc);
  if constexpr (G > 1) {
    static_for<0, RB>([&](auto ib) {
      static_for<0, NIMG>([&](auto jb) {
        acc[decltype(ib)::value][decltype(jb)::value] = group_sum<G>(acc[decltype(ib)::value][decltype(jb)::value]);
      });
    });
  }
  const bool owner = valid && (G == 1 || (j % G) == 0);
  // detector row of acc[ib]: lane (row group j) holds rows k = j*R .. j*R+R-1
  auto det_row = [&](int ib) { return R >= B ? (F - j * R) / B - 1 - ib : (F - 1 - j * R) / B; };

  // ---- detector: normalise, sky, noise, threshold, power, CoG window sums ----
  const int M2 = NIMG * NIMG;
  const long pix0 = (long)s * M2;
  const float scale = valid ? a.scale[s] : 1.f;
  const float skyadd = valid ? a.skyadd[s] : 0.f;
  const int st0 = (NIMG - a.ncen) / 2, en0 = NIMG - st0;
  float sw = 0.f, sxw = 0.f, syw = 0.f;
  if (!a.slow_detector) {
    // no noise, sky array, relative floor, power or weights: all inline, the threshold mode and
    // the image store decided once per item rather than per pixel
    const float add = (skyadd > 0.f ? skyadd : 0.f) + (a.threshType >= 0 && a.threshType <= 2 ? a.readbg : 0.f);
    const float thr = a.floor;
    const int tt = a.threshType;
    const bool full = st0 == 0;   // the CoG window is the whole image
    // Every pixel is a sum of |X|^2 times a positive scale, so a zero floor (type 0) or a floor
    // <= 0 (type 1) changes none. With nothing added either and no image returned, the CoG is
    // that of the bins: the scale cancels in the ratio, and the sums are taken on the bins
    // (for one detector row per lane; with more, the extra path costs registers).
    const bool thr_noop = tt < 0 || (tt == 0 && thr == 0.f) || (tt == 1 && thr <= 0.f);
    if (RB == 1 && full && !a.bimg && add == 0.f && thr_noop && scale > 0.f) {
      static_for<0, RB>([&](auto ibb) {
        constexpr int ib = decltype(ibb)::value;
        float rs = 0.f, rx = 0.f;
        static_for<0, NIMG>([&](auto jbb) {
          constexpr int jb = decltype(jbb)::value;
          rs += acc[ib][jb];
          rx = fmaf((float)jb, acc[ib][jb], rx);
        });
        sw += rs;
        sxw += rx;
        syw = fmaf((float)det_row(ib), rs, syw);   // the lane's row shares one detector row index
      });
    } else {
      // per detector row: pixel values (the threshold mode by uniform selects), the image row
      // stored under one owner test, CoG sums
      const bool store = a.bimg && owner;
      float *bout = a.bimg + (long)frame * a.bimg_stride + pix0;
      static_for<0, RB>([&](auto ibb) {
        constexpr int ib = decltype(ibb)::value;
        const int I = det_row(ib);
        static_for<0, NIMG>([&](auto jbb) {
          float &v = acc[ib][decltype(jbb)::value];
          v = fmaf(v, scale, 0.f) + add;
          const float v0 = fmaxf(v - thr, 0.f);   // type 0: (v < thr ? 0 : v - thr)
          const float v1 = v < thr ? 0.f : v;     // type 1
          v = tt == 0 ? v0 : (tt == 1 ? v1 : v);
        });
        if (store) {
          static_for<0, NIMG>([&](auto jbb) { bout[I * NIMG + decltype(jbb)::value] = acc[ib][decltype(jbb)::value]; });
        }
        float rs = 0.f, rx = 0.f;
        static_for<0, NIMG>([&](auto jbb) {
          constexpr int jb = decltype(jbb)::value;
          const float v = (full || (jb >= st0 && jb < en0)) ? acc[ib][jb] : 0.f;
          rs += v;
          rx = fmaf((float)jb, v, rx);
        });
        if (full || (I >= st0 && I < en0)) {
          sw += rs;
          sxw += rx;
          syw = fmaf((float)I, rs, syw);
        }
      });
    }
    if (!owner) sw = sxw = syw = 0.f;
  } else {
    // general detector: stage the lanes' pixels in the wave's LDS region and walk them with
    // runtime loops, so that the rarely used per-pixel code is emitted once
    float *px = reinterpret_cast<float *>(wl) + (sl * NP + j) * PXS;
    static_for<0, RB>([&](auto ib) {
      static_for<0, NIMG>([&](auto jb) {
        px[decltype(ib)::value * NIMG + decltype(jb)::value] = acc[decltype(ib)::value][decltype(jb)::value];
      });
    });
    // the loops' uniform arguments in registers (no argument-segment reload per pixel)
    const long noff = (long)frame * a.noise_stride + pix0;
    const bool pois = valid && a.poisson_on && a.poisson_on[s];
    const float *skyarr = a.skyarr ? a.skyarr + pix0 : nullptr;
    const float *zpois = a.zpois + noff, *zread = a.zread ? a.zread + noff : nullptr;
    const float readnoise = a.readnoise, readbg = a.readbg, pxlPower = a.pxlPower;
    const int threshType = a.threshType;
    const bool rn = readnoise != 0.f && zread;
    const bool addbg = threshType >= 0 && threshType <= 2;
    const float sky = skyadd > 0.f ? skyadd : 0.f;
    const float bg = addbg ? readbg : 0.f;
    const bool owner_pois = owner && pois;
    const bool rn_bg = rn && addbg;   // read noise only where readCCD adds the background
    float mx = -INFINITY;
    // four pixels per trip (consecutive pixels of one detector row): their pre-drawn normals are
    // two 16-byte loads, issued one trip ahead so that a trip's memory latency overlaps the
    // previous trip's work; the pointers are null only when unused
    const bool havez = a.zpois != nullptr, haver = zread != nullptr;
    auto pixel = [&](int q) { return det_row(q / NIMG) * NIMG + q % NIMG; };
    static_assert(PX % 4 == 0 && NIMG % 4 == 0, "a lane's pixels come in row-aligned fours");
    auto loadz = [&](int q0, float4 &zp, float4 &zr) {
      const int p = pixel(q0);   // pixels p .. p + 3, 16-byte aligned (noff and p are multiples of 4)
      zp = havez ? *(const float4 *)(zpois + p) : make_float4(0.f, 0.f, 0.f, 0.f);
      zr = haver ? *(const float4 *)(zread + p) : make_float4(0.f, 0.f, 0.f, 0.f);
    };
    float4 zpn, zrn;
    loadz(0, zpn, zrn);
#pragma unroll 1
    for (int q0 = 0; q0 < PX; q0 += 4) {
      const float zp4[4] = {zpn.x, zpn.y, zpn.z, zpn.w}, zr4[4] = {zrn.x, zrn.y, zrn.z, zrn.w};
      loadz(q0 + 4 < PX ? q0 + 4 : q0, zpn, zrn);   // the last trip reloads its own pixels
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int q = q0 + t;
        float v = px[q] * scale;   // scale is 1 where the theoretical total is not positive
        if (owner) {
          const int p = pixel(q);
          if (skyadd > 0.f) v += skyadd;
          if (skyarr) v += skyarr[p];
          if (pois) v = poisson_from_normal(v, zp4[t]);
          if (addbg) v += rn ? fmaf(readnoise, zr4[t], readbg) : readbg;
          mx = fmaxf(mx, v);
        }
        px[q] = v;
      }
    }
    float thr = a.floor;
    if (threshType == 2) {
      mx = group_max<NP>(mx);
      if (a.floor > 0.f && a.floor <= 1.f) thr = a.floor * fmaxf(mx, 0.f);
    }
    int tt = threshType;   // the threshold the pixel loop below applies
    {
      // brightest-pixel selection (cmod/src/centmodule.c:1544-1581), after readCCD's threshold:
      // the ub-th largest of the M2 values (with repeats) is the cut, found by a bitwise search
      // on order-preserving integer keys that counts the group's values at or above each
      // candidate; the selection is group-uniform (one item per lane group)
      int ub = valid ? (a.useBrightestArr ? a.useBrightestArr[s] : a.useBrightest) : 0;
      const bool bsub = ub < 0;
      if (bsub) ub = -ub;
      if (ub > 0 && ub < M2) {
        // keys in registers up to 16 pixels a lane; above that in place of the values in LDS
        // (the registers would lower the occupancy of the kernel as a whole), turned back into
        // values when the cut is applied
        constexpr bool KREG = PX <= 16;
        unsigned kreg[KREG ? PX : 1];
        unsigned *klds = reinterpret_cast<unsigned *>(px);
        auto key_of = [](float v) {
          const unsigned u = __builtin_bit_cast(unsigned, v);
          return (u >> 31) ? ~u : (u | 0x80000000u);
        };
        auto value_of = [](unsigned k) { return __builtin_bit_cast(float, (k >> 31) ? (k & 0x7fffffffu) : ~k); };
        static_for<0, PX>([&](auto qq) {
          constexpr int q = decltype(qq)::value;
          float v = px[q];
          if (tt == 1) v = v < thr ? 0.f : v;
          else if (tt == 0 || tt == 2) v = v < thr ? 0.f : v - thr;
          v = v == 0.f ? 0.f : v;   // -0 as +0: keys then order and tie as the values compare
          // key 0 (below every value's key) for lanes that hold no pixels of their own
          const unsigned k = owner ? key_of(v) : 0u;
          if constexpr (KREG) {
            kreg[q] = k;
            px[q] = v;
          } else {
            klds[q] = k;
          }
        });
        auto key = [&](auto qq) -> unsigned {
          if constexpr (KREG) return kreg[decltype(qq)::value];
          else return klds[decltype(qq)::value];
        };
        tt = -1;
        unsigned cut = 0u;
        const float need = (float)ub;
#pragma unroll 1
        for (int b = 31; b >= 0; --b) {
          const unsigned cand = cut | (1u << b);
          int cnt = 0;
          static_for<0, PX>([&](auto qq) { cnt += key(qq) >= cand ? 1 : 0; });
          if (group_sum<NP>((float)cnt) >= need) cut = cand;   // counts below 2^24: exact
        }
        const float cv = value_of(cut);
        float sub = 0.f;
        if (bsub) {
          // the largest value below the cut, 0 when there is none
          float lo = -INFINITY;
          static_for<0, PX>([&](auto qq) {
            const unsigned k = key(qq);
            if (k != 0u && k < cut) lo = fmaxf(lo, value_of(k));
          });
          lo = group_max<NP>(lo);
          sub = lo == -INFINITY ? 0.f : lo;
        }
        static_for<0, PX>([&](auto qq) {
          constexpr int q = decltype(qq)::value;
          float v;
          if constexpr (KREG) v = px[q];
          else v = value_of(klds[q]);
          px[q] = v < cv ? 0.f : v - sub;
        });
      }
    }
    float *bout = a.bimg ? a.bimg + (long)frame * a.bimg_stride + pix0 : nullptr;
    const float *wt = a.centWeight ? a.centWeight + (a.centWeightDim == 4 ? pix0 : 0) : nullptr;
    if (pxlPower == 1.f && !wt) {
      // no pixel power and no weights (the usual noisy detector): the threshold by selects and
      // the window by a per-lane row test and compile-time columns, unrolled
      const bool t1 = tt == 1, t02 = tt == 0 || tt == 2;
      const float tsub = t02 ? thr : 0.f;
      const bool thr_on = t1 || t02;
      static_for<0, PX>([&](auto qq) {
        constexpr int q = decltype(qq)::value, jb = q % NIMG;
        const int I = det_row(q / NIMG);
        float v = px[q];
        v = (thr_on && v < thr) ? 0.f : v - tsub;
        if (bout && owner) bout[I * NIMG + jb] = v;
        const bool in = owner && I >= st0 && I < en0 && jb >= st0 && jb < en0;
        const float t = in ? v : 0.f;
        sw += t;
        sxw = fmaf(t, (float)jb, sxw);
        syw = fmaf(t, (float)I, syw);
      });
    } else {
#pragma unroll 1
    for (int q = 0; q < PX; ++q) {
      const int I = det_row(q / NIMG), jb = q % NIMG;
      float v = px[q];
      if (tt == 1) v = v < thr ? 0.f : v;
      else if (tt == 0 || tt == 2) v = v < thr ? 0.f : v - thr;
      if (pxlPower == 2.f) v = v * v;
      else if (pxlPower != 1.f && pxlPower != 0.f) v = powf(v, pxlPower);
      if (owner) {
        if (bout) bout[I * NIMG + jb] = v;
        if (I >= st0 && I < en0 && jb >= st0 && jb < en0) {
          const float t = wt ? v * wt[I * NIMG + jb] : v;
          sw += t;
          sxw += t * jb;
          syw += t * I;
        }
      }
    }
    }
    wave_lds_sync();   // the region holds the next item's transposes
  }
  sw = group_sum<NP>(sw);
  sxw = group_sum<NP>(sxw);
  syw = group_sum<NP>(syw);
  if (valid && j == 0 && !a.no_cents) {
    float cx = 0.f, cy = 0.f;
    if (sw > 0.f) {
      cx = sxw / sw - (NIMG - 1) / 2.f;
      cy = syw / sw - (NIMG - 1) / 2.f;
    }
    if (a.refCents) {
      cx -= a.refCents[s * 2];
      cy -= a.refCents[s * 2 + 1];
    }
    float *co = a.cents + (long)frame * a.cents_stride + s * 2;
    co[0] = cx;
    co[1] = cy;
  }
}

}  // namespace dasp
