// cent_dft32.hip - the DFT-GEMM form for phase size n <= 32, fftsize 64, nimg 16 (4x4 binning).
//
// One sub-aperture per wave. computeHll's zero-padded 2-D transform (cmod/src/centmodule.c:
// 473-509) as two GEMMs with W[k][i] = exp(-2 pi i k i / 64), on v_mfma_f32_32x32x16_{f16,bf16}
// with the three-product hi/lo scheme of dft_ops.h:
//   stage 1: Z^T = E^T W^T   M = 32 (j), N = 64 (k: two tiles kt), K = 32 (i: two steps ks)
//   stage 2: X^T = W Z^T     M = 64 (l: two tiles lt), N = 64 (k: two tiles kt), K = 32 (j)
// 48 + 96 = 144 MFMA per image. Stage 1's accumulator of tile kt holds Z^T[j][k] with lane
// k - 32 kt + 32 h and register r <-> j = 8(r/4) + 4h + r%4, so registers 8 ks .. 8 ks + 7 are
// the eight K slots of stage 2's B operand for K step ks when W's K index is permuted the same
// way (slot q <-> j = 16 ks + 8(q/4) + 4h + q%4): no transpose. Stage 2's accumulator of tile
// (lt, kt) holds X[k][l] with lane = k - 32 kt + 32 h and register 4g + u <-> l = 32 lt + 8g + 4h
// + u: a 4x4 detector pixel is four registers in-lane (u) summed over the four lanes of a quad
// (k), and after the 180-degree map it is pixel (I, J) = (15 - 8 kt - c/4, 15 - 8 lt - 2g - h).
// The DFT-matrix operands (eight of them) are built once per workgroup into LDS. COMP selects the
// three-product scheme; without it one product of the FP16 roundings (the single-product form).
#include "dft_ops.h"
#include "noise.h"
#include "wave_ops.h"

namespace dasp {

namespace {
constexpr int kF64 = 64, kM64 = 16, kXcds = 8;
}


template <typename T8, bool FAST, bool FULL, bool COMP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(COMP || !FULL ? 2 : 3))) void dft32_cog_kernel(FusedArgs a_) {
  typedef const __attribute__((address_space(4))) FusedArgs KernArgs;
  const FusedArgs &a = *(const FusedArgs *)(KernArgs *)__builtin_amdgcn_kernarg_segment_ptr();
  constexpr int F = kF64, M = kM64;
  __shared__ CConstLds<T8, COMP> w1[2][2];   // [kt][ks]: W[32 kt + c][16 ks + 8h + t]
  __shared__ CConstLds<T8, COMP> w2[2][2];   // [lt][ks]: W[32 lt + c][16 ks + 8(t/4) + 4h + t%4]
  __shared__ float pxbuf[4][64 * 4];
  const int l = threadIdx.x & 63;
  const int h = l >> 5;
  const int c = l & 31;
  float *px = &pxbuf[threadIdx.x >> 6][l * 4];
  if (threadIdx.x < 64) {
#pragma unroll 1
    for (int tile = 0; tile < 2; ++tile)
#pragma unroll 1
      for (int ks = 0; ks < 2; ++ks) {
        float re[8], im[8], re2[8], im2[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const int i = 16 * ks + 8 * h + t;
          const int k = 32 * tile + c;
          sincospif(-2.0f * (float)((k * i) % F) / F, &im[t], &re[t]);
          const int j = 16 * ks + 8 * (t / 4) + 4 * h + t % 4;
          sincospif(-2.0f * (float)((k * j) % F) / F, &im2[t], &re2[t]);
        }
        CConst<T8, COMP> v1, v2;
        v1.set(re, im);
        v2.set(re2, im2);
        w1[tile][ks].put(l, v1);
        w2[tile][ks].put(l, v2);
      }
  }
  __syncthreads();
  // XCD-aware item order: workgroup b runs on XCD b % 8, and XCD x takes the x-th eighth of the
  // active sub-apertures for every frame. Within an XCD the slice's items are taken with a grid
  // stride in sub-aperture-major order (all frames of one sub-aperture, then the next), so that the
  // XCD's waves work on one or two sub-apertures at a time and their pupils stay in L2 (in
  // frame-major order the slice's whole pupil, 4 KB per sub-aperture at n = 32, competed with the
  // streamed phase and was partly read again from HBM for every frame).
  const int xcd = blockIdx.x % kXcds;
  const int s0 = (int)((long)a.nactive * xcd / kXcds);
  const int cnt = (int)((long)a.nactive * (xcd + 1) / kXcds) - s0;
  const long nx = cnt > 0 ? (a.nitems / a.nactive) * cnt : 0;   // this XCD's items
  const long wave_id = (long)(blockIdx.x / kXcds) * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long nwaves = (long)(gridDim.x / kXcds) * (blockDim.x >> 6);
  zero_inactive_slopes(a);
  if (wave_id >= nx) return;   // wave-uniform; the workgroup has no barrier after this point
  const int n = a.n;
  // local item q = r * nfr + f (sub-aperture s0 + r of frame f), q advancing by nwaves: one
  // scalar division per wave, then a fixed step
  const int nfr = (int)(a.nitems / a.nactive);
  auto split_item = [&](long q, int &f, int &r) {
    const long rq = q / nfr;
    r = (int)rq;
    f = (int)(q - rq * nfr);
  };
  int stepF, stepR;
  split_item(nwaves, stepF, stepR);
  // lane (c, h) holds phase column j = c, rows i = 16 ks + 8h + t
  // Loads are unconditional (no per-element branches): past the last item the wave reads item 0
  // (never used), and for n < 32 the padding rows/columns read a clamped in-window address with
  // the pupil factor forced to 0, so E is exactly 0 there.
  auto load_item = [&](long q, int f, int r, float (&ph)[16], float (&pu)[16]) {
    const bool v = q < nx;   // wave-uniform
    const int s = v ? scalar_read(a.active, s0 + r) : 0;
    const float *p0 = a.phs + (v ? (long)f * a.phs_stride + scalar_read(a.winoff, s0 + r) : 0);
    const float *u0 = a.pup + (long)s * n * n;
    if constexpr (FULL) {
      // wave-uniform row bases (scalar registers) plus one 32-bit lane byte offset
      unsigned po = (unsigned)(8 * h * a.npup + c) * 4u, uo = (unsigned)(8 * h * 32 + c) * 4u;
      pin_offset(po);
      pin_offset(uo);
      GlobalBytes *ub = uniform_global(u0);
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int r = 16 * (q >> 3) + (q & 7);
        ph[q] = load_at(uniform_global(p0 + (long)r * a.npup), po);
        pu[q] = load_at(ub + r * 128, uo);
      }
    } else {
      // the zero factor is a product, not a select, so that the load is not made conditional
      const int cc = c < n ? c : n - 1;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int i = 16 * (q >> 3) + 8 * h + (q & 7);
        const int ii = i < n ? i : n - 1;
        ph[q] = p0[ii * a.npup + cc];
        pu[q] = u0[ii * n + cc] * ((i < n && c < n) ? 1.f : 0.f);
      }
    }
  };
  int cf, cidx;
  split_item(wave_id, cf, cidx);
  float cph[16], cpu[16];
  load_item(wave_id, cf, cidx, cph, cpu);
  for (long q = wave_id; q < nx; q += nwaves) {
    int nf = cf + stepF, nidx = cidx + stepR;
    if (nf >= nfr) {
      nf -= nfr;
      ++nidx;
    }
    const int frame = cf, s = scalar_read(a.active, s0 + cidx);
    // ---- E = pup exp(i (phi - t)) and its hi/lo operands ----
    COp<T8, COMP> A1[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      float er[8], ei[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int i = 16 * ks + 8 * h + t;
        const float th = cph[8 * ks + t] - a.tiltstep * (float)(i + c + 1 - n);   // float ops as the reference
        float sn, cs;
        dft_sincos<FAST>(th, &sn, &cs);
        er[t] = cpu[8 * ks + t] * cs;
        ei[t] = cpu[8 * ks + t] * sn;
      }
      A1[ks].set(er, ei);
    }
    // reload the LDS constants where used (an opaque zero offset keeps the reads in the loop)
    int lofs = 0;
    asm volatile("" : "+v"(lofs));
    // ---- per k tile: stage 1 (Z^T tile, split into stage 2's B operands), then stage 2 for
    // both l tiles with |X|^2 binned 4x4: in-lane over u, over the quad's four lanes along k ----
    float pix[2][2];   // [lt][kt]: this lane's pixel of the tile
    const bool b0 = (c & 1) != 0, b1 = (c & 2) != 0;
#pragma unroll
    for (int kt = 0; kt < 2; ++kt) {
      COp<T8, COMP> B2[2];   // [ks]
      {
        f16v zr = {}, zi = {};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
          const CConst<T8, COMP> w = w1[kt][ks].get(l + lofs);
          cgemm_vc(A1[ks], w, zr, zi);
        }
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
          float br[8], bi[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            br[q] = zr[8 * ks + q];
            bi[q] = zi[8 * ks + q];
          }
          B2[ks].set(br, bi);
        }
      }
      // the next item's loads, issued once the stage-1 operands are dead (their registers are
      // not held across the whole item) and overlapping the stage-2 products
      if (kt == 1) load_item(q + nwaves, nf, nidx, cph, cpu);
#pragma unroll
      for (int lt = 0; lt < 2; ++lt) {
        f16v xr = {}, xi = {};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
          const CConst<T8, COMP> w = w2[lt][ks].get(l + lofs);
          cgemm_cv(w, B2[ks], xr, xi);
        }
        float v[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const int r = 4 * g;
          float t = fmaf(xr[r], xr[r], xi[r] * xi[r]);
          t = fmaf(xr[r + 1], xr[r + 1], fmaf(xi[r + 1], xi[r + 1], t));
          t = fmaf(xr[r + 2], xr[r + 2], fmaf(xi[r + 2], xi[r + 2], t));
          v[g] = fmaf(xr[r + 3], xr[r + 3], fmaf(xi[r + 3], xi[r + 3], t));
        }
        // reduce-scatter over the quad (its four lanes are four k of one detector row): quad
        // lane p ends with the total of pixel column group g = p. Step 1 exchanges with lane
        // p ^ 2 and keeps g in {2 b1, 2 b1 + 1}; step 2 exchanges with p ^ 1 and keeps g = p.
        const float k0 = b1 ? v[2] : v[0], k1 = b1 ? v[3] : v[1];
        const float s0 = b1 ? v[0] : v[2], s1 = b1 ? v[1] : v[3];
        const float w0 = k0 + dpp_mov<dpp::kQuadXor2>(s0);
        const float w1 = k1 + dpp_mov<dpp::kQuadXor2>(s1);
        pix[lt][kt] = (b0 ? w1 : w0) + dpp_mov<dpp::kQuadXor1>(b0 ? w0 : w1);
      }
    }
    // ---- detector: lane (c, h) takes g = c % 4 of every tile: four distinct pixels ----
    const int g = c & 3;
    float val[4];
    int pI[4], pJ[4];
#pragma unroll
    for (int lt = 0; lt < 2; ++lt)
#pragma unroll
      for (int kt = 0; kt < 2; ++kt) {
        const int e = 2 * lt + kt;
        val[e] = pix[lt][kt];
        pI[e] = M - 1 - 8 * kt - (c >> 2);
        pJ[e] = M - 1 - 8 * lt - 2 * g - h;
      }
    const int M2 = M * M;
    const long pix0 = (long)s * M2;
    const float scale = scalar_read(a.scale, s);
    const float skyadd = scalar_read(a.skyadd, s);
    const int st0 = (M - a.ncen) / 2, en0 = M - st0;
    cf = nf;
    cidx = nidx;
    float sw = 0.f, sxw = 0.f, syw = 0.f;
    if (!a.slow_detector) {
      const float add = (skyadd > 0.f ? skyadd : 0.f) + (a.threshType >= 0 && a.threshType <= 2 ? a.readbg : 0.f);
      const float thr = a.floor;
      float *bout = a.bimg ? a.bimg + (long)frame * a.bimg_stride + pix0 : nullptr;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        float v = fmaf(val[e], scale, 0.f) + add;
        if (a.threshType == 1) v = v < thr ? 0.f : v;
        else if (a.threshType == 0) v = v < thr ? 0.f : v - thr;
        if (bout) bout[pI[e] * M + pJ[e]] = v;
        if (pI[e] >= st0 && pI[e] < en0 && pJ[e] >= st0 && pJ[e] < en0) {
          sw += v;
          sxw += v * pJ[e];
          syw += v * pI[e];
        }
      }
    } else {
      // general detector in an LDS-staged runtime loop (rare options, code emitted once)
#pragma unroll
      for (int e = 0; e < 4; ++e) px[e] = val[e];
      const long noff = (long)frame * a.noise_stride + pix0;
      const bool pois = a.poisson_on && a.poisson_on[s];
      const bool rn = a.readnoise != 0.f && a.zread;
      float mx = -INFINITY;
#pragma unroll 1
      for (int e = 0; e < 4; ++e) {
        const int I = M - 1 - 8 * (e & 1) - (c >> 2), J = M - 1 - 8 * (e >> 1) - 2 * g - h;
        const long p = I * M + J;
        float v = px[e] * scale;
        if (skyadd > 0.f) v += skyadd;
        if (a.skyarr) v += a.skyarr[pix0 + p];
        if (pois) v = poisson_from_normal(v, a.zpois[noff + p]);
        if (a.threshType >= 0 && a.threshType <= 2) v += rn ? fmaf(a.readnoise, a.zread[noff + p], a.readbg) : a.readbg;
        mx = fmaxf(mx, v);
        px[e] = v;
      }
      float thr = a.floor;
      if (a.threshType == 2) {
        mx = wave_maximum(mx);
        if (a.floor > 0.f && a.floor <= 1.f) thr = a.floor * fmaxf(mx, 0.f);
      }
      float *bout = a.bimg ? a.bimg + (long)frame * a.bimg_stride + pix0 : nullptr;
      const float *wt = a.centWeight ? a.centWeight + (a.centWeightDim == 4 ? pix0 : 0) : nullptr;
#pragma unroll 1
      for (int e = 0; e < 4; ++e) {
        const int I = M - 1 - 8 * (e & 1) - (c >> 2), J = M - 1 - 8 * (e >> 1) - 2 * g - h;
        float v = px[e];
        if (a.threshType == 1) v = v < thr ? 0.f : v;
        else if (a.threshType == 0 || a.threshType == 2) v = v < thr ? 0.f : v - thr;
        if (a.pxlPower == 2.f) v = v * v;
        else if (a.pxlPower != 1.f && a.pxlPower != 0.f) v = powf(v, a.pxlPower);
        if (bout) bout[I * M + J] = v;
        if (I >= st0 && I < en0 && J >= st0 && J < en0) {
          const float tt = wt ? v * wt[I * M + J] : v;
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
        const float rw = __builtin_amdgcn_rcpf(sw);   // 1 ulp, inside this form's hi/lo error
        cx = sxw * rw - (M - 1) / 2.f;
        cy = syw * rw - (M - 1) / 2.f;
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

bool dft32_supported(int n, int F, int nimg) { return n > 16 && n <= 32 && F == kF64 && nimg == kM64; }

// Resident workgroups per CU of one kernel instance (occupancy query, once per instance).
template <typename T8, bool FAST, bool FULL, bool COMP>
static long dft32_resident() {
  static int per_cu = 0, cus = 0;
  if (per_cu == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, dft32_cog_kernel<T8, FAST, FULL, COMP>, 256, 0) !=
            hipSuccess || per_cu < 1 || cus < 1) {
      per_cu = 2;
      cus = 256;
    }
  }
  return (long)per_cu * cus;
}

template <typename T8, bool FAST, bool FULL, bool COMP>
static void launch_dft32_f(const FusedArgs &a, long blocks, hipStream_t st, long grid) {
  // resident workgroups, then grid-stride (the compensated kernels hold two waves per SIMD; three
  // waves of them, at 168 VGPRs, measured no faster), or the tuned grid
  blocks = launch_grid(blocks, dft32_resident<T8, FAST, FULL, COMP>(), grid);
  hipLaunchKernelGGL((dft32_cog_kernel<T8, FAST, FULL, COMP>), dim3((unsigned)blocks), dim3(256), 0, st, a);
}

template <typename T8, bool FAST, bool COMP>
static void launch_dft32_t(const FusedArgs &a, long blocks, hipStream_t st, long grid) {
  if (a.n == 32) launch_dft32_f<T8, FAST, true, COMP>(a, blocks, st, grid);
  else launch_dft32_f<T8, FAST, false, COMP>(a, blocks, st, grid);
}

template <typename T8, bool COMP>
static void launch_dft32_c(const FusedArgs &a, long blocks, hipStream_t st, long grid) {
  if (a.fast_trig) launch_dft32_t<T8, true, COMP>(a, blocks, st, grid);
  else launch_dft32_t<T8, false, COMP>(a, blocks, st, grid);
}

hipError_t launch_dft32(const FusedArgs &a, int bf16, bool comp, hipStream_t st, long grid) {
  if (a.nitems <= 0) return hipSuccess;
  const long blocks = (a.nitems + 3) / 4;
  if (bf16) {
    launch_dft32_c<b8, true>(a, blocks, st, grid);   // the single-product form is FP16 only
  } else {
    if (comp) launch_dft32_c<h8, true>(a, blocks, st, grid);
    else launch_dft32_c<h8, false>(a, blocks, st, grid);
  }
  return hipGetLastError();
}

}  // namespace dasp
