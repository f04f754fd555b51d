// cent_dft.hip - the DFT-GEMM form on the gfx950 matrix cores.
//
// The zero-padded 2-D transform of computeHll (cmod/src/centmodule.c:473-509) for n = 16,
// F = 32 written as two small GEMMs with the DFT matrix W[k][i] = exp(-2 pi i k i / F):
//   stage 1 (two subapertures stacked):  Z^T = E^T W^T      (M = 2x16 (s, j), K = 16 (i), N = 32 (k))
//   stage 2 (per subaperture):           X^T = W  Z^T       (M = 32 (l), K = 16 (j), N = 32 (k))
// Each complex product is four real ones and each real product is computed from a hi/lo split of
// both operands into 16-bit floats, hi*hi + hi*lo + lo*hi, accumulated in FP32 by
// v_mfma_f32_32x32x16_{f16,bf16} (a three-product compensated scheme). W is built in FP32 from
// exact sincospi values.
//
// MFMA 32x32x16 lane layout (checked on the device): A[m][kk] sits in lane m + 32*(kk/8), slot kk%8;
// B[kk][nn] in lane nn + 32*(kk/8), slot kk%8; D[m][nn] in lane nn + 32*((m/4)%2), register
// 4*(m/8) + m%4. Stage 1's accumulator therefore holds Z^T[j][k] with lane = k and, for lane half
// h, j in {4h..4h+3, 8+4h..8+4h+3}: read as a B operand of stage 2 those eight registers are
// eight K slots, so stage 2 uses W with its K index permuted the same way and needs no transpose.
// The accumulator of stage 2 holds X[k][l] with lane = k and l = 8g + 4h + u (g, u < 4): |X|^2
// is binned in-lane along l and across lane pairs along k, then the detector steps and the CoG
// follow as in the fused FP32 kernel.
// Two kernels: dft16_pipe_kernel (software pipelined, inline detector only) and dft16_cog_kernel
// (every detector option, including noise and relative floors).
#include "dft_ops.h"
#include "noise.h"
#include "wave_ops.h"

namespace dasp {

constexpr int kN = 16, kF = 32, kM = 16;   // phase size, fft size, detector size (bin 2)
constexpr int kXcds = 8;

// COMP: the three-product scheme; without it the single-product FP16 form (as in the pipelined
// kernel), which here carries the general detector options (noise, sky arrays, relative floors).
// The single-product form loads each pair when the wave starts it and runs three waves per SIMD:
// prefetching the next pair's operands held 16 VGPRs more and kept it at two, and the third wave
// hides the loads better (40x40 noisy workload 0.178 -> 0.165 ms, same-box A/B). The three-product
// forms, at two waves either way, keep the prefetch.
template <typename T8, bool FAST, bool COMP = true>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(COMP ? 2 : 3))) void dft16_cog_kernel(FusedArgs a_) {
  // arguments by value (in scalar registers; some spill to vector-register lanes): reading them
  // through the argument-segment pointer, as the FP32 kernels do, measured 1-2.5 % slower here
  const FusedArgs &a = a_;
  __shared__ float pxbuf[4][64 * 8];
  const int l = threadIdx.x & 63;
  const int h = l >> 5;
  const int c = l & 31;
  float *px = &pxbuf[threadIdx.x >> 6][l * 8];
  // constant operands: B1 = W^T (B1[i][k] = W[k][i], lane holds k = c, i = 8h + t) and
  // A2 = W with permuted K (A2[l][slot q] = W[l][j], j = 8(q/4) + 4h + q%4, lane holds l = c)
  CConst<T8, COMP> B1, A2;
  {
    float re[8], im[8], re2[8], im2[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int i = 8 * h + t;
      sincospif(-2.0f * (float)((c * i) % kF) / kF, &im[t], &re[t]);
      const int j = 8 * (t / 4) + 4 * h + t % 4;
      sincospif(-2.0f * (float)((c * j) % kF) / kF, &im2[t], &re2[t]);
    }
    B1.set(re, im);
    A2.set(re2, im2);
  }
  // XCD-aware pair order: workgroup b runs on XCD b % 8, and XCD x takes the x-th eighth of the
  // active sub-apertures for every frame (its L2 keeps that slice's pupil). A pair is two
  // neighbouring sub-apertures of the slice in one frame (the slice's last pair of a frame may
  // hold one); the XCD's pairs are taken frame-major with a grid stride.
  const int xcd = blockIdx.x % kXcds;
  const int s0 = (int)((long)a.nactive * xcd / kXcds);
  const int cnt = (int)((long)a.nactive * (xcd + 1) / kXcds) - s0;
  const int pp = (cnt + 1) / 2;                                   // pairs per frame
  const long npairs = cnt > 0 ? (a.nitems / a.nactive) * pp : 0;  // this XCD's pairs
  const long wave_id = (long)(blockIdx.x / kXcds) * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long nwaves = (long)(gridDim.x / kXcds) * (blockDim.x >> 6);
  zero_inactive_slopes(a);
  if (wave_id >= npairs) return;   // wave-uniform; no barrier follows
  const int n = a.n;
  const int jj = c & 15;
  // fast-detector constants of this lane: its detector row (lane pair k = c, c+1: hll rows F-1-k,
  // F-2-k), its four pixel columns (lane parity p takes q = 4p..4p+3; l pair starts at
  // l0 = 8(q/2) + 4h + 2(q%2), J = (F-2-l0)/2) and their CoG weights inside the ncen window
  int pJ[4];
  float wc[4], wj[4];
  float fI;
  {
    const int st0 = (kM - a.ncen) / 2, en0 = kM - st0;
    const int I = (kF - 2 - (c & ~1)) >> 1;
    fI = (float)I;
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) {
      const int q = 4 * (c & 1) + qq;
      const int J = (kF - 2 - (8 * (q >> 1) + 4 * h + 2 * (q & 1))) >> 1;
      const bool in = I >= st0 && I < en0 && J >= st0 && J < en0;
      pJ[qq] = J;
      wc[qq] = in ? 1.f : 0.f;
      wj[qq] = in ? (float)J : 0.f;
    }
  }
  // frame and index within the slice of a pair (the pair index is wave-uniform: one scalar
  // division per wave, then a fixed step)
  auto pair_ref = [&](long pr, int &f, int &p) {
    const long q = pr / pp;
    f = (int)q;
    p = (int)(pr - q * pp);
  };
  // stage-1 operand of a pair: lane holds E_s[i][j], s = c/16, j = c%16, i = 8h + t
  auto load_pair = [&](long pr, int f, int p, float (&ph8)[8], float (&pu8)[8]) {
    const int u = c >> 4;
    // both items' sub-aperture indices and window offsets in scalar registers (scalar-cache
    // reads, scalar division), then the lane's item by a select. Loads are unconditional: an
    // item past the end reads item 0 (its results are never used).
    const bool vA = pr < npairs, vB = vA && 2 * p + 1 < cnt;
    const int fA = f, fB = f;
    const int sA = vA ? scalar_read(a.active, s0 + 2 * p) : 0;
    const int sB = vB ? scalar_read(a.active, s0 + 2 * p + 1) : 0;
    auto window = [&](int sa, int fr) {
      const int sy = sa / a.nsubx, sx = sa - sy * a.nsubx;
      return (long)fr * a.phs_stride + (long)(sy * kN) * a.npup + sx * kN;
    };
    const long wA = window(sA, vA ? fA : 0), wB = window(sB, vB ? fB : 0);
    const float *ph = a.phs + (u ? wB : wA) + jj;
    const float *pu = a.pup + (long)(u ? sB : sA) * (kN * kN) + jj;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int i = 8 * h + t;
      ph8[t] = ph[(long)i * a.npup];
      pu8[t] = pu[i * kN];
    }
  };
  float cph[8], cpu[8];
  int cf0, cidx0;
  pair_ref(wave_id, cf0, cidx0);
  // the pair index advances by nwaves: frame / active index advance by a fixed step (no division
  // inside the loop)
  int stepF, stepR;
  pair_ref(nwaves, stepF, stepR);
  if constexpr (COMP) load_pair(wave_id, cf0, cidx0, cph, cpu);
  for (long pr = wave_id; pr < npairs; pr += nwaves) {
    float nph[8], npu[8];
    int nf0 = cf0 + stepF, nidx0 = cidx0 + stepR;
    if (nidx0 >= pp) {
      nidx0 -= pp;
      ++nf0;
    }
    if constexpr (COMP) load_pair(pr + nwaves, nf0, nidx0, nph, npu);   // overlaps this pair's work
    else load_pair(pr, cf0, cidx0, cph, cpu);
    float er[8], ei[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int i = 8 * h + t;
      const float th = cph[t] - a.tiltstep * (float)(i + jj + 1 - n);   // float ops as the reference
      float sn, cs;
      dft_sincos<FAST>(th, &sn, &cs);
      er[t] = cpu[t] * cs;
      ei[t] = cpu[t] * sn;
    }
    if constexpr (COMP) {
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        cph[t] = nph[t];
        cpu[t] = npu[t];
      }
    }
    const int pf0 = cf0, pidx0 = cidx0;
    cf0 = nf0;
    cidx0 = nidx0;
    COp<T8, COMP> A1;
    A1.set(er, ei);
    f16v zr = {}, zi = {};
    cgemm_vc(A1, B1, zr, zi);   // Z^T stacked: rows (s, j), columns k

#pragma unroll
    for (int sub = 0; sub < 2; ++sub) {
      if (2 * pidx0 + sub >= cnt) break;   // wave-uniform
      const int frame = pf0;
      const int s = scalar_read(a.active, s0 + 2 * pidx0 + sub);
      // ---- stage 2: B2 = Z_s^T from accumulator registers 8*sub .. 8*sub+7 ----
      float br[8], bi[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        br[q] = zr[8 * sub + q];
        bi[q] = zi[8 * sub + q];
      }
      COp<T8, COMP> B2;
      B2.set(br, bi);
      f16v xr = {}, xi = {};
      cgemm_cv(A2, B2, xr, xi);   // X^T: rows l, columns k (lane)
      // ---- |X|^2 binned 2x2: in-lane along l, lane pairs along k ----
      // register r holds l = 8(r/4) + 4h + r%4; pairs (r, r+1) for even r share a detector column
      float pix[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int r = 2 * q;
        float v = fmaf(xr[r], xr[r], fmaf(xi[r], xi[r], fmaf(xr[r + 1], xr[r + 1], xi[r + 1] * xi[r + 1])));
        v += dpp_mov<dpp::kQuadXor1>(v);   // the lane pair (k, k^1) of one detector row
        pix[q] = v;
      }
      // both lanes of a pair hold the pair's sums: lane parity p takes pixel columns q = 4p..4p+3
      const int par = c & 1;
      // detector row of the lane pair (k = c, c+1): hll rows F-1-k, F-2-k -> I = (F-2-(c & ~1))/2
      const int I = (kF - 2 - (c & ~1)) >> 1;
      const int M2 = kM * kM;
      const long pix0 = (long)s * M2;
      const float scale = scalar_read(a.scale, s);
      const float skyadd = scalar_read(a.skyadd, s);
      const int st0 = (kM - a.ncen) / 2, en0 = kM - st0;
      float sw = 0.f, sxw = 0.f, syw = 0.f;
      if (!a.slow_detector) {
        const float add = (skyadd > 0.f ? skyadd : 0.f) + (a.threshType >= 0 && a.threshType <= 2 ? a.readbg : 0.f);
        const float thr = a.floor;
        float *bout = a.bimg ? a.bimg + (long)frame * a.bimg_stride + pix0 : nullptr;
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) {
          float v = fmaf(par ? pix[4 + qq] : pix[qq], scale, 0.f) + add;
          if (a.threshType == 1) v = v < thr ? 0.f : v;
          else if (a.threshType == 0) v = fmaxf(v - thr, 0.f);   // = (v < thr ? 0 : v - thr)
          if (bout) bout[I * kM + pJ[qq]] = v;
          sw = fmaf(wc[qq], v, sw);
          sxw = fmaf(wj[qq], v, sxw);
        }
        syw = fI * sw;   // the lane's pixels share one detector row
      } else {
        // general detector: each lane of the pair takes four of its eight pixel sums (columns
        // q = 4p .. 4p+3 as above), so that every lane draws noise for distinct pixels
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) px[qq] = par ? pix[4 + qq] : pix[qq];
        const long noff = (long)frame * a.noise_stride + pix0;
        const bool pois = a.poisson_on && a.poisson_on[s];
        const bool rn = a.readnoise != 0.f && a.zread;
        // the lane's columns are J0, J0-1, J0-4, J0-5 (J0 odd): its pre-drawn normals are two
        // 8-byte-aligned pairs per array, loaded before the pixels are walked
        const int J0 = pJ[0];
        float zp[4] = {0.f, 0.f, 0.f, 0.f}, zr[4] = {0.f, 0.f, 0.f, 0.f};
        if (pois) {
          const float2 z0 = *(const float2 *)(a.zpois + noff + I * kM + J0 - 1);
          const float2 z1 = *(const float2 *)(a.zpois + noff + I * kM + J0 - 5);
          zp[0] = z0.y; zp[1] = z0.x; zp[2] = z1.y; zp[3] = z1.x;
        }
        if (rn) {
          const float2 z0 = *(const float2 *)(a.zread + noff + I * kM + J0 - 1);
          const float2 z1 = *(const float2 *)(a.zread + noff + I * kM + J0 - 5);
          zr[0] = z0.y; zr[1] = z0.x; zr[2] = z1.y; zr[3] = z1.x;
        }
        float mx = -INFINITY;
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) {
          const int J = J0 - (qq == 0 ? 0 : (qq == 1 ? 1 : (qq == 2 ? 4 : 5)));
          float v = px[qq] * scale;
          if (skyadd > 0.f) v += skyadd;
          if (a.skyarr) v += a.skyarr[pix0 + I * kM + J];
          if (pois) v = poisson_from_normal(v, zp[qq]);
          if (a.threshType >= 0 && a.threshType <= 2) v += rn ? fmaf(a.readnoise, zr[qq], a.readbg) : a.readbg;
          mx = fmaxf(mx, v);
          px[qq] = v;
        }
        float thr = a.floor;
        if (a.threshType == 2) {
          mx = wave_maximum(mx);
          if (a.floor > 0.f && a.floor <= 1.f) thr = a.floor * fmaxf(mx, 0.f);
        }
        float *bout = a.bimg ? a.bimg + (long)frame * a.bimg_stride + pix0 : nullptr;
        const float *wt = a.centWeight ? a.centWeight + (a.centWeightDim == 4 ? pix0 : 0) : nullptr;
        if (a.pxlPower == 1.f && !wt) {
          // no pixel power and no weights (the usual noisy detector): selects, unrolled
          const bool t1 = a.threshType == 1, t02 = a.threshType == 0 || a.threshType == 2;
          const bool thr_on = t1 || t02;
          const float tsub = t02 ? thr : 0.f;
#pragma unroll
          for (int qq = 0; qq < 4; ++qq) {
            const int J = J0 - (qq == 0 ? 0 : (qq == 1 ? 1 : (qq == 2 ? 4 : 5)));
            float v = px[qq];
            v = (thr_on && v < thr) ? 0.f : v - tsub;
            if (bout) bout[I * kM + J] = v;
            const float tt = (I >= st0 && I < en0 && J >= st0 && J < en0) ? v : 0.f;
            sw += tt;
            sxw = fmaf(tt, (float)J, sxw);
            syw = fmaf(tt, (float)I, syw);
          }
        } else
#pragma unroll 1
        for (int qq = 0; qq < 4; ++qq) {
          const int J = pJ[0] - (qq == 0 ? 0 : (qq == 1 ? 1 : (qq == 2 ? 4 : 5)));
          float v = px[qq];
          if (a.threshType == 1) v = v < thr ? 0.f : v;
          else if (a.threshType == 0 || a.threshType == 2) v = v < thr ? 0.f : v - thr;
          if (a.pxlPower == 2.f) v = v * v;
          else if (a.pxlPower != 1.f && a.pxlPower != 0.f) v = powf(v, a.pxlPower);
          if (bout) bout[I * kM + J] = v;
          if (I >= st0 && I < en0 && J >= st0 && J < en0) {
            const float tt = wt ? v * wt[I * kM + J] : v;
            sw += tt;
            sxw += tt * J;
            syw += tt * I;
          }
        }
      }
      sw = wave_total(sw);
      sxw = wave_total(sxw);
      syw = wave_total(syw);
      if (l == 0 && !a.no_cents) {
        float cx = 0.f, cy = 0.f;
        if (sw > 0.f) {
          // hardware reciprocal (1 ulp) instead of two IEEE divisions: a few ulp on the slopes,
          // far inside this form's hi/lo error
          const float rw = __builtin_amdgcn_rcpf(sw);
          cx = sxw * rw - (kM - 1) / 2.f;
          cy = syw * rw - (kM - 1) / 2.f;
        }
        if (a.refCents) {
          cx -= scalar_read(a.refCents, s * 2);
          cy -= scalar_read(a.refCents, s * 2 + 1);
        }
        float *co = a.cents + (long)frame * a.cents_stride + s * 2;
        co[0] = cx;
        co[1] = cy;
      }
    }
  }
}

// Software-pipelined form for the inline detector (no noise, sky array, relative floor, pixel
// power or weights): one loop iteration runs stage 2 of pair t and stage 1 of pair t+1, with the
// phase factor of pair t+1 and the detector of pair t as the vector work beside those products.
// Each iteration is one basic block up to its stores (thresholds by selects, the second item of a
// pair computed even when the slice has none and then not stored), so the scheduler interleaves
// the matrix-core instructions with independent vector instructions of the same wave.
//
// The half-pixel tilt is folded into the constant operands: exp(-i t) with
// t = tiltstep*(i + j + 1 - n) factors into exp(-i tiltstep*i) on stage 1's contraction index and
// exp(-i tiltstep*(j + 1 - n)) on stage 2's, so the phase factor is pup*exp(i phi) and the tilt
// costs no vector instruction or register per pixel (the products are formed in FP64 before the
// hi/lo split; the reference subtracts the float tilt from the phase, which rounds the argument).
// FULLWIN (ncen == nimg): the CoG window is the whole detector, so no per-pixel window weights.
// Register budget: the FP16 kernel with fast trig and the full window fits three waves per SIMD
// (163 VGPRs) with one wave_total per sum; the transposing six-value sum saves 40 vector
// instructions per pair but needs 171 VGPRs (two waves). Measured on the 80x80 nfft32 workload,
// the first is 1-2 % faster for FP16 and the second 2-3 % faster for BF16 (whose split code makes
// it a two-wave kernel either way), so each type takes its faster variant.
template <typename T8, bool FAST, bool FULLWIN, bool COMP>
constexpr bool kPipeThreeWaves = COMP && std::is_same<T8, h8>::value && FAST && FULLWIN;
// waves per SIMD asked of the compiler (1: its own choice)
template <typename T8, bool FAST, bool FULLWIN, bool COMP>
constexpr int kPipeWaves = !COMP ? (FAST ? 3 : 2) : (kPipeThreeWaves<T8, FAST, FULLWIN, COMP> ? 3 : 1);

// COMP: the three-product scheme; without it the single-product form (one product of the FP16
// roundings, a third of the matrix-core work and no lo parts). BINS (with FULLWIN, when the
// host's bins_only holds): the CoG is taken on the binned |X|^2 itself, which the positive
// scale, the absent additions and the no-op threshold leave unchanged up to a factor.
template <typename T8, bool FAST, bool FULLWIN, bool COMP, bool BINS = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kPipeWaves<T8, FAST, FULLWIN, COMP>)))
void dft16_pipe_kernel(FusedArgs a_) {
  static_assert(!BINS || FULLWIN, "the bins-only detector takes the whole image as window");
  const FusedArgs &a = a_;
  const int l = threadIdx.x & 63;
  const int h = l >> 5;
  const int c = l & 31;
  const int jj = c & 15;
  CConst<T8, COMP> B1, A2;   // as in dft16_cog_kernel, times the tilt factors
  {
    float re[8], im[8], re2[8], im2[8];
    const double ts = (double)a.tiltstep, twopi_f = 6.283185307179586 / kF;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int i = 8 * h + t;
      double sn, cs;
      sincos(-twopi_f * (double)((c * i) % kF) - ts * (double)i, &sn, &cs);
      re[t] = (float)cs;
      im[t] = (float)sn;
      const int j = 8 * (t / 4) + 4 * h + t % 4;
      sincos(-twopi_f * (double)((c * j) % kF) - ts * (double)(j + 1 - a.n), &sn, &cs);
      re2[t] = (float)cs;
      im2[t] = (float)sn;
    }
    B1.set(re, im);
    A2.set(re2, im2);
  }
  const int xcd = blockIdx.x % kXcds;
  const int s0 = (int)((long)a.nactive * xcd / kXcds);
  const int cnt = (int)((long)a.nactive * (xcd + 1) / kXcds) - s0;
  const int pp = (cnt + 1) / 2;
  const long npairs = cnt > 0 ? (a.nitems / a.nactive) * pp : 0;
  const long wave_id = (long)(blockIdx.x / kXcds) * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long nwaves = (long)(gridDim.x / kXcds) * (blockDim.x >> 6);
  zero_inactive_slopes(a);
  if (wave_id >= npairs) return;   // wave-uniform; no barrier in this kernel
  // detector constants of this lane (see dft16_cog_kernel): row I, and columns J0 - d with
  // d = 0, 1, 4, 5 for its four pixels
  const int I = (kF - 2 - (c & ~1)) >> 1;
  const float fI = (float)I;
  const int par = c & 1;
  const int J0 = kM - 1 - 8 * par - 2 * h;
  const float fJ0 = (float)J0;
  constexpr int kD[4] = {0, 1, 4, 5};
  float wc[4], wj[4];
  if constexpr (!FULLWIN) {
    const int st0 = (kM - a.ncen) / 2, en0 = kM - st0;
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) {
      const int J = J0 - kD[qq];
      const bool in = I >= st0 && I < en0 && J >= st0 && J < en0;
      wc[qq] = in ? 1.f : 0.f;
      wj[qq] = in ? (float)J : 0.f;
    }
  }
  // thresholds as selects: v < thr ? 0 : v - tsub for types 0 and 1, no threshold otherwise
  const bool thr_on = a.threshType == 0 || a.threshType == 1;
  const float thr = a.floor, tsub = a.threshType == 0 ? a.floor : 0.f;
  const float readbg = a.threshType >= 0 && a.threshType <= 2 ? a.readbg : 0.f;
  auto pair_ref = [&](long pr, int &f, int &p) {
    const long q = pr / pp;
    f = (int)q;
    p = (int)(pr - q * pp);
  };
  // a pair's sub-aperture indices and window offsets (scalar reads), fetched one iteration before
  // its loads are issued so that the loads do not wait for them
  struct PairIdx {
    int rA, rB, oA, oB;
  };
  auto fetch_idx = [&](int p) {
    // unconditional in-slice reads (p < pp), selected afterwards: no branch in the loop body;
    // window offsets from the per-sub-aperture table (no division)
    const int iA = s0 + 2 * p, iB = s0 + min(2 * p + 1, cnt - 1);
    return PairIdx{scalar_read(a.active, iA), scalar_read(a.active, iB), scalar_read(a.winoff, iA),
                   scalar_read(a.winoff, iB)};
  };
  auto load_pair = [&](long pr, int f, int p, const PairIdx &ix, float (&ph8)[8], float (&pu8)[8]) {
    const int u = c >> 4;
    const bool vA = pr < npairs, vB = vA && 2 * p + 1 < cnt;
    const int rA = ix.rA, rB = ix.rB, oA = ix.oA, oB = ix.oB;
    // Item A's phase window and pupil are wave-uniform bases; the lanes of item B add B's
    // distance from A (non-negative: active indices and window offsets ascend), so each load is
    // a scalar base plus one 32-bit lane offset, with no 64-bit address arithmetic per load.
    const long wA = vA ? (long)f * a.phs_stride + oA : 0;
    const int sA = vA ? rA : 0;
    const unsigned dW = vB ? (unsigned)(oB - oA) : 0u, dS = vB ? (unsigned)(rB - rA) : 0u;
    const unsigned po = ((u ? dW : 0u) + (unsigned)(8 * h * a.npup + jj)) * 4u;
    const unsigned uo = ((u ? dS * (kN * kN) : 0u) + (unsigned)(8 * h * kN + jj)) * 4u;
    const char *ph = (const char *)(a.phs + wA);
    const char *pu = (const char *)(a.pup + (long)sA * (kN * kN));
    const long rstride = (long)a.npup * 4;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      ph8[t] = load_at(uniform_global(ph + t * rstride), po);
      pu8[t] = *(const float *)(pu + t * (kN * 4) + uo);
    }
  };
  auto phase_factor = [&](const float (&ph8)[8], const float (&pu8)[8], COp<T8, COMP> &A1) {
    float er[8], ei[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      float sn, cs;
      if constexpr (FAST && !COMP) {
        // single-product form: the fraction of a revolution straight from one product (its
        // rounding, at most 2.4e-5 rad at |phi| = 400 rad, is far inside the FP16 rounding of
        // the operands, 4.9e-4 relative), then the hardware sin/cos of revolutions
        const float r = __builtin_amdgcn_fractf(ph8[t] * 0.15915494309189535f);
        sn = __builtin_amdgcn_sinf(r);
        cs = __builtin_amdgcn_cosf(r);
      } else {
        dft_sincos<FAST>(ph8[t], &sn, &cs);   // the tilt is in B1 and A2
      }
      er[t] = pu8[t] * cs;
      ei[t] = pu8[t] * sn;
    }
    A1.set(er, ei);
  };
  int stepF, stepR;
  pair_ref(nwaves, stepF, stepR);
  auto advance = [&](int &f, int &p) {
    f += stepF;
    p += stepR;
    if (p >= pp) {
      p -= pp;
      ++f;
    }
  };
  // prologue: stage 1 of the first pair, loads of the second
  int cf, cp;
  pair_ref(wave_id, cf, cp);
  float ph8[8], pu8[8];
  load_pair(wave_id, cf, cp, fetch_idx(cp), ph8, pu8);
  f16v zr = {}, zi = {};
  {
    COp<T8, COMP> A1;
    phase_factor(ph8, pu8, A1);
    cgemm_vc(A1, B1, zr, zi);
  }
  int nf = cf, np = cp;
  advance(nf, np);
  load_pair(wave_id + nwaves, nf, np, fetch_idx(np), ph8, pu8);
  int xf = nf, xp = np;   // the pair two ahead of the current one: its indices are fetched early
  advance(xf, xp);
  PairIdx xidx = fetch_idx(xp);
  for (long q = wave_id; q < npairs; q += nwaves) {
    // Three scheduling regions (nothing crosses a sched_barrier), each pairing matrix-core work
    // with independent vector work: (1) stage 2 of pair q with the phase factor of the pair
    // after it, (2) stage 1 of that pair with the detector of item 0 of pair q, (3) the detector
    // of item 1, the wave sums and the loads two pairs ahead.
    const int f = cf, p = cp;
    // ---- region 1 ----
    COp<T8, COMP> B2[2];
#pragma unroll
    for (int sub = 0; sub < 2; ++sub) {
      float br[8], bi[8];
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        br[r] = zr[8 * sub + r];
        bi[r] = zi[8 * sub + r];
      }
      B2[sub].set(br, bi);
    }
    f16v xr0 = {}, xi0 = {}, xr1 = {}, xi1 = {};
    cgemm_cv(A2, B2[0], xr0, xi0);
    cgemm_cv(A2, B2[1], xr1, xi1);
    COp<T8, COMP> A1n;
    phase_factor(ph8, pu8, A1n);   // pair q + nwaves (loaded one iteration ago)
    __builtin_amdgcn_sched_barrier(0);
    // ---- region 2 ----
    zr = f16v{};
    zi = f16v{};
    cgemm_vc(A1n, B1, zr, zi);
    // detector of pair q: |X|^2 binned 2x2, normalisation, threshold, CoG sums
    float sums[2][3];
    float img[2][4];
    int sidx[2];
    auto detector = [&](int sub, const f16v &xr, const f16v &xi) {
      float pix[8];
#pragma unroll
      for (int r8 = 0; r8 < 8; ++r8) {
        const int r = 2 * r8;
        float v = fmaf(xr[r], xr[r], fmaf(xi[r], xi[r], fmaf(xr[r + 1], xr[r + 1], xi[r + 1] * xi[r + 1])));
        v += dpp_mov<dpp::kQuadXor1>(v);
        pix[r8] = v;
      }
      const bool valid = 2 * p + sub < cnt;
      const int s = scalar_read(a.active, s0 + 2 * p + (valid ? sub : 0));
      sidx[sub] = s;
      if constexpr (BINS) {
        float v[4];
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) v[qq] = par ? pix[4 + qq] : pix[qq];
        const float sw = (v[0] + v[1]) + (v[2] + v[3]);
        sums[sub][0] = sw;
        sums[sub][1] = fmaf(fJ0, sw, -fmaf(5.f, v[3], fmaf(4.f, v[2], v[1])));
        sums[sub][2] = fI * sw;
        return;
      }
      const float scale = scalar_read(a.scale, s);
      const float skyadd = scalar_read(a.skyadd, s);
      const float add = (skyadd > 0.f ? skyadd : 0.f) + readbg;
      float sw = 0.f, sxw = 0.f;
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) {
        float v = fmaf(par ? pix[4 + qq] : pix[qq], scale, 0.f) + add;
        v = (thr_on && v < thr) ? 0.f : v - tsub;
        img[sub][qq] = v;
        if constexpr (!FULLWIN) {
          sw = fmaf(wc[qq], v, sw);
          sxw = fmaf(wj[qq], v, sxw);
        }
      }
      if constexpr (FULLWIN) {
        // sum of v * (J0 - d) = J0 * sum(v) - (v1 + 4 v2 + 5 v3)
        const float(&v)[4] = img[sub];
        sw = (v[0] + v[1]) + (v[2] + v[3]);
        sxw = fmaf(fJ0, sw, -fmaf(5.f, v[3], fmaf(4.f, v[2], v[1])));
      }
      sums[sub][0] = sw;
      sums[sub][1] = sxw;
      sums[sub][2] = fI * sw;   // the lane's pixels share one detector row
    };
    detector(0, xr0, xi0);
    __builtin_amdgcn_sched_barrier(0);
    // ---- region 3 ----
    detector(1, xr1, xi1);
    if constexpr (kPipeThreeWaves<T8, FAST, FULLWIN, COMP>) {
#pragma unroll
      for (int sub = 0; sub < 2; ++sub)
#pragma unroll
        for (int k = 0; k < 3; ++k) sums[sub][k] = wave_total(sums[sub][k]);
    } else {
      wave_totals6(sums[0], sums[1]);
    }
    cf = nf;
    cp = np;
    advance(nf, np);
    load_pair(q + 2 * nwaves, nf, np, xidx, ph8, pu8);   // two pairs ahead
    advance(xf, xp);
    xidx = fetch_idx(xp);                                  // three pairs ahead
    // ---- stores ----
    const int nvalid = 2 * p + 1 < cnt ? 2 : 1;
    if (!BINS && a.bimg) {   // bins_only holds only without an image buffer
#pragma unroll
      for (int sub = 0; sub < 2; ++sub) {
        if (sub < nvalid) {
          char *bout = (char *)(a.bimg + (long)f * a.bimg_stride + (long)sidx[sub] * (kM * kM)) +
                       (unsigned)((I * kM + J0) * 4);
#pragma unroll
          for (int qq = 0; qq < 4; ++qq) *(float *)(bout - 4 * kD[qq]) = img[sub][qq];
        }
      }
    }
    if (l == 0 && !a.no_cents) {
#pragma unroll
      for (int sub = 0; sub < 2; ++sub) {
        if (sub < nvalid) {
          const int s = sidx[sub];
          float cx = 0.f, cy = 0.f;
          if (sums[sub][0] > 0.f) {
            const float rw = __builtin_amdgcn_rcpf(sums[sub][0]);
            cx = sums[sub][1] * rw - (kM - 1) / 2.f;
            cy = sums[sub][2] * rw - (kM - 1) / 2.f;
          }
          if (a.refCents) {
            cx -= scalar_read(a.refCents, s * 2);
            cy -= scalar_read(a.refCents, s * 2 + 1);
          }
          float *co = a.cents + (long)f * a.cents_stride + s * 2;
          co[0] = cx;
          co[1] = cy;
        }
      }
    }
  }
}

// Persistent launch of the pipelined kernel: every wave builds the DFT-matrix operands once, so
// the grid is the resident wave count (workgroups per CU from the occupancy query, times the CU
// count) or the pair count if smaller, in whole multiples of the XCD count.
template <typename T8, bool FAST, bool FULLWIN, bool COMP, bool BINS = false>
hipError_t launch_pipe_t(const FusedArgs &a, long npairs, hipStream_t st, long grid) {
  static int per_cu = 0, cus = 0;
  if (per_cu == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, dft16_pipe_kernel<T8, FAST, FULLWIN, COMP, BINS>, 256, 0) !=
            hipSuccess || per_cu < 1 || cus < 1) {
      per_cu = 2;
      cus = 256;
    }
  }
  const long blocks = launch_grid((npairs + 3) / 4, (long)cus * per_cu, grid);
  hipLaunchKernelGGL((dft16_pipe_kernel<T8, FAST, FULLWIN, COMP, BINS>), dim3((unsigned)blocks), dim3(256), 0, st, a);
  return hipGetLastError();
}

template <typename T8, bool FAST, bool COMP>
hipError_t launch_pipe_c(const FusedArgs &a, bool full, long npairs, hipStream_t st, long grid) {
  if (full && a.bins_only) return launch_pipe_t<T8, FAST, true, COMP, true>(a, npairs, st, grid);
  return full ? launch_pipe_t<T8, FAST, true, COMP>(a, npairs, st, grid)
              : launch_pipe_t<T8, FAST, false, COMP>(a, npairs, st, grid);
}

template <typename T8, bool FAST>
hipError_t launch_pipe(const FusedArgs &a, bool full, long npairs, hipStream_t st, long grid) {
  return launch_pipe_c<T8, FAST, true>(a, full, npairs, st, grid);
}

bool dft_supported(int n, int F, int nimg) {
  return (n == kN && F == kF && nimg == kM) || dft32_supported(n, F, nimg);
}

hipError_t launch_dft(const FusedArgs &a, int bf16, bool comp, hipStream_t st, long grid) {
  if (a.nitems <= 0) return hipSuccess;
  if (a.n != kN) return launch_dft32(a, bf16, comp, st, grid);
  const long npairs = (a.nitems + 1) / 2 + kXcds;
  // persistent grid for the general-detector kernel: every wave builds the DFT-matrix operands
  // once (hundreds of instructions), so launch its resident wave count (two waves per SIMD)
  const long blocks = launch_grid((npairs + 3) / 4, 256L * 2, grid);
  if (!a.slow_detector) {
    const bool full = a.ncen == kM;
    if (bf16) {
      if (a.fast_trig) return launch_pipe<b8, true>(a, full, npairs, st, grid);
      return launch_pipe<b8, false>(a, full, npairs, st, grid);
    }
    if (!comp) {   // the single-product form is FP16 only
      if (a.fast_trig) return launch_pipe_c<h8, true, false>(a, full, npairs, st, grid);
      return launch_pipe_c<h8, false, false>(a, full, npairs, st, grid);
    }
    if (a.fast_trig) return launch_pipe<h8, true>(a, full, npairs, st, grid);
    return launch_pipe<h8, false>(a, full, npairs, st, grid);
  }
  // the general detector options: the unpipelined kernel, three products or (FP16) one
  if (!comp && !bf16) {
    if (a.fast_trig) hipLaunchKernelGGL((dft16_cog_kernel<h8, true, false>), dim3((unsigned)blocks), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((dft16_cog_kernel<h8, false, false>), dim3((unsigned)blocks), dim3(256), 0, st, a);
    return hipGetLastError();
  }
  if (bf16) {
    if (a.fast_trig) hipLaunchKernelGGL((dft16_cog_kernel<b8, true>), dim3((unsigned)blocks), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((dft16_cog_kernel<b8, false>), dim3((unsigned)blocks), dim3(256), 0, st, a);
  } else {
    if (a.fast_trig) hipLaunchKernelGGL((dft16_cog_kernel<h8, true>), dim3((unsigned)blocks), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((dft16_cog_kernel<h8, false>), dim3((unsigned)blocks), dim3(256), 0, st, a);
  }
  return hipGetLastError();
}

}  // namespace dasp
