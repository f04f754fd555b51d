// cent_backend.hip - per-subaperture image pipeline after image formation, general form.
//
// One workgroup (one wave) per (frame, subaperture). The image lives in LDS; every option of
// centroidsFromPhase / centroidsFromImage (cmod/src/centmodule.c:1808-1961, :2000-2110) is
// supported: pre-binning, LGS spot convolution, fractional binning, normalisation, sky,
// optical binning, Poisson and read noise from pre-drawn normals, floor thresholds, brightest
// pixel selection, pixel power, FFT / difference / brute correlation, CoG (plain, weighted,
// centred on the brightest or faintest pixel), parabolic / Gaussian / 3x3 fits, linearisation
// calibration and reference slopes. The fused kernel (cent_fused.hip) is the fast path for the
// common configurations; this one is the general path.
#include "cent_internal.h"
#include "noise.h"
#include "wave_ops.h"

namespace dasp {

constexpr int BT = 64;             // threads per workgroup: one wave per item

// The workgroup's LDS, carved from dynamic shared memory sized for the handle's geometry (a
// 32 x 32 image with a 16 x 16 detector needs about 11 KB, where room for the largest supported
// side, 64, took 66 KB and held the kernel to two workgroups per CU).
struct BackLds {
  float *h;                        // hll / convolved image / correlation image
  float2 *c;                       // complex work buffer; also sort and binning scratch
  float *img;                      // detector image (nimg x nimg)
  float *red;                      // BT
  int *ired;                       // BT
  float2 *tw;                      // kMaxSide / 2
};

// Sizes (elements) of the regions for one launch: h floats, c complex values, img floats.
struct BackLdsSizes {
  int h, c, img;
  __host__ __device__ BackLdsSizes(const BackArgs &a) {
    auto mx = [](int x, int y) { return x > y ? x : y; };
    auto sq = [](int x) { return x * x; };
    const int M2 = sq(a.nimg);
    int n2 = 1;
    while (n2 < M2) n2 <<= 1;
    h = mx(mx(sq(a.padded), sq(a.psf)), mx(mx(sq(a.corrsize), M2), sq(a.clip)));
    // the complex buffer only for the options that use it: spot convolution, FFT correlation,
    // pre-binning, the separable (fractional) binning and the brightest-pixel sort
    const bool sort = a.useBrightest != 0 || a.useBrightestArr != nullptr;
    c = mx(mx(a.otf ? sq(a.psf) : 0, a.corrMode == 1 ? sq(a.corrsize) : 0),
           (mx(mx(a.pbf > 1 ? sq(a.padded) : 0, (a.binB > 0 || a.clip == a.nimg) ? 0 : a.nimg * a.clip),
               sort ? n2 : 0) + 1) / 2);
    img = M2;
  }
  __host__ __device__ size_t bytes() const {
    return (size_t)c * sizeof(float2) + (size_t)(kMaxSide / 2) * sizeof(float2) + (size_t)h * sizeof(float) +
           (size_t)img * sizeof(float) + (size_t)BT * (sizeof(float) + sizeof(int));
  }
};

// Sums and maxima over the workgroup (one wave): cross-lane (DPP) reductions, the result in every
// lane. Called by every lane of the wave.
__device__ __forceinline__ float block_sum(float v, BackLds &) {
  static_assert(BT == 64, "one wave per workgroup");
  return group_sum<64>(v);
}

__device__ __forceinline__ float block_max(float v, BackLds &) { return group_max<64>(v); }

__device__ __forceinline__ int bitrev_n(int v, int bits) {
  return (int)(__brev((unsigned)v) >> (32 - bits));
}

// In-place 2-D FFT of an N x N complex array in LDS (N a power of two <= 64).
// sign = -1: forward (FFTW_FORWARD); sign = +1: unnormalised inverse.
__device__ void lds_fft2(float2 *buf, int N, int sign, BackLds &L) {
  int bits = 0;
  while ((1 << bits) < N) ++bits;
  for (int q = threadIdx.x; q < N / 2; q += BT) {
    float sn, cs;
    sincospif(-2.0f * q / N, &sn, &cs);
    L.tw[q] = make_float2(cs, sign > 0 ? -sn : sn);
  }
  for (int pass = 0; pass < 2; ++pass) {
    // element (r, c) of the line layout: pass 0 transforms rows, pass 1 columns
    for (int idx = threadIdx.x; idx < N * N; idx += BT) {
      int r = idx / N, c = idx - (idx / N) * N;
      int rc = bitrev_n(c, bits);
      if (c < rc) {
        int a = pass == 0 ? r * N + c : c * N + r;
        int b = pass == 0 ? r * N + rc : rc * N + r;
        float2 t = buf[a];
        buf[a] = buf[b];
        buf[b] = t;
      }
    }
    __syncthreads();
    for (int h = 1; h < N; h <<= 1) {
      for (int b = threadIdx.x; b < N * N / 2; b += BT) {
        int r = b / (N / 2);
        int t = b - r * (N / 2);
        int g = t / h;
        int jj = t - g * h;
        int i0 = g * 2 * h + jj, i1 = i0 + h;
        float2 w = L.tw[jj * (N / (2 * h))];
        int a0 = pass == 0 ? r * N + i0 : i0 * N + r;
        int a1 = pass == 0 ? r * N + i1 : i1 * N + r;
        float2 u = buf[a0], v = buf[a1];
        float2 vw = make_float2(fmaf(v.x, w.x, -v.y * w.y), fmaf(v.x, w.y, v.y * w.x));
        buf[a0] = make_float2(u.x + vw.x, u.y + vw.y);
        buf[a1] = make_float2(u.x - vw.x, u.y - vw.y);
      }
      __syncthreads();
    }
  }
}

// Bitonic sort (ascending) of n2 (power of two) floats in LDS.
__device__ void lds_sort(float *v, int n2) {
  for (int k = 2; k <= n2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < n2; i += BT) {
        int ixj = i ^ j;
        if (ixj > i) {
          bool up = (i & k) == 0;
          float a = v[i], b = v[ixj];
          if ((a > b) == up) {
            v[i] = b;
            v[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
}

// computeCoG (cmod/src/centmodule.c:873-971) on an N x N LDS image.
__device__ void cog(const float *img, int N, int ncen, const float *wt, int cob, float *cx, float *cy, BackLds &L) {
  int sx, sy;
  if (cob == 0) {
    sx = (N - ncen) / 2;
    sy = sx;
  } else {
    // serial scan: the reference's strict comparisons decide ties (first max; faintest rule)
    if (threadIdx.x == 0) {
      float mx = 0.f;
      int pos = 0;
      if (cob == 1) {
        for (int i = 0; i < N * N; ++i)
          if (img[i] > mx) { mx = img[i]; pos = i; }
      } else {
        for (int i = 0; i < N * N; ++i)
          if (img[i] < mx || mx == 0) { mx = img[i]; pos = i; }
      }
      int ax = (int)(pos % N - ncen / 2. + 0.5);
      int ay = (int)(pos / N - ncen / 2. + 0.5);
      if (ax < 0) ax = 0;
      if (ay < 0) ay = 0;
      if (ax > N - ncen) ax = N - ncen;
      if (ay > N - ncen) ay = N - ncen;
      L.ired[0] = ax;
      L.ired[1] = ay;
    }
    __syncthreads();
    sx = L.ired[0];
    sy = L.ired[1];
    __syncthreads();
  }
  float s = 0.f, ax = 0.f, ay = 0.f;
  const bool neg = (wt == nullptr) && !(cob == 0 || cob == 1);
  for (int q = threadIdx.x; q < ncen * ncen; q += BT) {
    int i = sy + q / ncen, j = sx + q % ncen;
    float t = img[i * N + j];
    if (wt) t *= wt[i * N + j];
    s += t;
    ax += neg ? -t * j : t * j;
    ay += neg ? -t * i : t * i;
  }
  s = block_sum(s, L);
  ax = block_sum(ax, L);
  ay = block_sum(ay, L);
  if (s > 0) {
    ax /= s;
    ay /= s;
    if (cob != 0) {
      ax += sx;
      ay += sy;
    }
    *cx = ax - (N - 1) / 2.f;
    *cy = ay - (N - 1) / 2.f;
  } else {
    *cx = 0.f;
    *cy = 0.f;
  }
}

// computeFit (cmod/src/centmodule.c:1069-1188), serial on thread 0 (small windows).
__device__ void fit(float *img, int N, int ncen, int fitsize, const float *fitMx, int doGaussian,
                    float gmin, float grep, int corrMode, float *cx, float *cy) {
  const int nc = (N - ncen) / 2;
  bool quad = false;
  if (fitsize < 0) {
    quad = true;
    fitsize = 3;
  }
  if (fitsize < 3) fitsize = 3;
  float sum = 0.f, mx = 0.f;
  int mxpos = 0;
  if (doGaussian) {
    for (int i = 0; i < N * N; ++i) {
      sum += img[i];
      if (img[i] > mx) { mxpos = i; mx = img[i]; }
      img[i] = img[i] < gmin ? grep : logf(img[i]);
    }
  } else if (corrMode == 0 || corrMode == 1 || corrMode == 5) {
    for (int i = 0; i < N * N; ++i) {
      sum += img[i];
      if (img[i] > mx) { mxpos = i; mx = img[i]; }
    }
  } else {
    mx = img[nc * N + nc];
    mxpos = nc * N + nc;
    for (int i = nc; i < nc + ncen; ++i)
      for (int j = nc; j < nc + ncen; ++j) {
        sum += img[i * N + j];
        if (img[i * N + j] < mx) { mxpos = i * N + j; mx = img[i * N + j]; }
      }
  }
  if (!(sum >= 0.f)) {
    *cx = 0.f;
    *cy = 0.f;
    return;
  }
  int startx = mxpos % N - fitsize / 2;
  int starty = mxpos / N - fitsize / 2;
  if (startx < nc) startx = nc;
  if (starty < nc) starty = nc;
  if (startx > nc + ncen - fitsize) startx = nc + ncen - fitsize;
  if (starty > nc + ncen - fitsize) starty = nc + ncen - fitsize;
  float x, y;
  if (!quad) {
    float vec[6] = {0, 0, 0, 0, 0, 0};
    for (int yy = 0; yy < fitsize; ++yy)
      for (int xx = 0; xx < fitsize; ++xx) {
        float val = img[(yy + starty) * N + xx + startx];
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
      for (int j = 0; j < 6; ++j) t += fitMx[i * 6 + j] * vec[j];
      res[i] = t;
    }
    if (doGaussian) {
      x = -res[3] / (2 * res[0]);
      y = -res[4] / (2 * res[2]);
    } else {
      x = (res[1] * res[4] / (2 * res[2]) - res[3]) / (2.f * res[0] - res[1] * res[1] / (2.f * res[2]));
      y = -(res[4] + res[1] * x) / (2.f * res[2]);
    }
  } else {
    const float *w = &img[starty * N + startx];
    float a2 = (w[2 * N + 1] - w[1]) / 2.f;
    float a3 = (w[2 * N + 1] - 2 * w[N + 1] + w[1]) / 2.f;
    float a4 = (w[N + 2] - w[N]) / 2.f;
    float a5 = (w[N + 2] - 2 * w[N + 1] + w[N]) / 2.f;
    float a6 = (w[2 * N + 2] - w[2] - w[2 * N] + w[0]) / 4.f;
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
  *cx = x;
  *cy = y;
}

__device__ float interp_cal(float x, int n, const float *xp, const float *fp) {
  if (x <= xp[0]) return fp[0];
  if (x >= xp[n]) return fp[n];
  int i = 1;
  while (x > xp[i]) ++i;
  return (x - xp[i - 1]) * (fp[i] - fp[i - 1]) / (xp[i] - xp[i - 1]) + fp[i - 1];
}

__global__ __launch_bounds__(BT) void backend_kernel(BackArgs a) {
  extern __shared__ float2 back_lds[];
  BackLds L;
  {
    const BackLdsSizes z(a);
    L.c = back_lds;
    L.tw = L.c + z.c;
    L.h = reinterpret_cast<float *>(L.tw + kMaxSide / 2);
    L.img = L.h + z.h;
    L.red = L.img + z.img;
    L.ired = reinterpret_cast<int *>(L.red + BT);
  }
  const int tid = threadIdx.x;
  const long item = blockIdx.x;
  const int frame = (int)(item / a.nactive);
  const int s = a.active[item - (long)frame * a.nactive];
  const int M = a.nimg, M2 = M * M;
  int nexposed = M2;
  float *img = L.img;

  if (a.inputImageDims == 0) {
    // ---------------- phase mode: hll -> (prebin) -> (spot) -> bin -> normalise ----------------
    int cur = a.padded;
    const float *src = a.hll + item * (long)cur * cur;
    for (int p = tid; p < cur * cur; p += BT) L.h[p] = src[p];
    __syncthreads();
    if (a.pbf > 1) {   // prebinImage (cmod/src/centmodule.c:1583-1626)
      const int scaled = cur / a.pbf;
      int binned = scaled, offset = 0;
      if (scaled > a.psf) {
        offset = (scaled - a.psf) / 2;
        binned = a.psf;
      }
      float *tmp = reinterpret_cast<float *>(L.c);
      for (int q = tid; q < binned * binned; q += BT) {
        int i = q / binned, j = q % binned;
        float t = 0.f;
        for (int ii = 0; ii < a.pbf; ++ii)
          for (int jj = 0; jj < a.pbf; ++jj)
            t += L.h[(i * a.pbf + ii + offset) * cur + j * a.pbf + jj + offset];
        tmp[q] = t;
      }
      __syncthreads();
      const int P = a.psf;
      const int off2 = scaled < P ? (P - scaled) / 2 : 0;
      for (int q = tid; q < P * P; q += BT) {
        int i = q / P - off2, j = q % P - off2;
        L.h[q] = (i >= 0 && j >= 0 && i < binned && j < binned) ? tmp[i * binned + j] : 0.f;
      }
      __syncthreads();
      cur = P;
    }
    if (a.otf) {   // doConvolution (cmod/src/centmodule.c:524-564)
      const int P = a.psf;
      for (int q = tid; q < P * P; q += BT) L.c[q] = make_float2(L.h[q], 0.f);
      __syncthreads();
      lds_fft2(L.c, P, -1, L);
      const float2 *K = a.otf + (a.otf_per_subap ? (long)s * P * P : 0);
      for (int q = tid; q < P * P; q += BT) {
        float2 u = L.c[q], k = K[q];
        L.c[q] = make_float2(fmaf(u.x, k.x, -u.y * k.y), fmaf(u.x, k.y, u.y * k.x));
      }
      __syncthreads();
      lds_fft2(L.c, P, +1, L);
      const int C = a.clip;
      const int off = (P - C) / 2;
      for (int q = tid; q < C * C; q += BT) L.h[q] = L.c[(q / C + off) * P + q % C + off].x;
      __syncthreads();
      cur = C;
    }
    // binImage through its separable operator A (nimg x clip): img = A h A^T
    const int Q = a.clip;
    if (Q == M) {
      for (int q = tid; q < M2; q += BT) img[q] = L.h[q];
    } else if (a.binB > 0) {
      // the plain block sum, in the operator's order: each column's B rows, then the B columns
      const int B = a.binB;
      for (int q = tid; q < M2; q += BT) {
        const int I = q / M, J = q % M;
        const float *hb = L.h + (I * B) * Q + J * B;
        float t = 0.f;
        for (int dj = 0; dj < B; ++dj) {
          float u = 0.f;
          for (int di = 0; di < B; ++di) u += hb[di * Q + dj];
          t += u;
        }
        img[q] = t;
      }
    } else {
      float *tmp = reinterpret_cast<float *>(L.c);
      for (int q = tid; q < M * Q; q += BT) {
        int I = q / Q, jcol = q % Q;
        float t = 0.f;
        for (int i = 0; i < Q; ++i) {
          float w = a.binA[I * Q + i];
          if (w != 0.f) t = fmaf(w, L.h[i * Q + jcol], t);
        }
        tmp[q] = t;
      }
      __syncthreads();
      for (int q = tid; q < M2; q += BT) {
        int I = q / M, J = q % M;
        float t = 0.f;
        for (int jcol = 0; jcol < Q; ++jcol) {
          float w = a.binA[J * Q + jcol];
          if (w != 0.f) t = fmaf(w, tmp[I * Q + jcol], t);
        }
        img[q] = t;
      }
    }
    __syncthreads();
    // normalisation (:1829-1863)
    float scale = a.scale[s];
    float totsig = a.totsig[s];
    if (a.amp) {   // phase+amplitude: total from the amplitude actually present
      const int n = a.n;
      const int sy = s / a.nsubx, sx = s % a.nsubx;
      const float *am = a.amp + (long)frame * a.phs_stride + (long)(sy * n) * a.npup + sx * n;
      const float *pu = a.pup + (long)s * n * n;
      float ts = 0.f, tp = 0.f;
      for (int q = tid; q < n * n; q += BT) {
        float p = pu[q];
        ts += p * am[(long)(q / n) * a.npup + q % n];
        tp += p;
      }
      ts = block_sum(ts, L);
      tp = block_sum(tp, L);
      totsig = tp != 0.f ? ts * (a.totsig[s] / tp) : 0.f;
      scale = totsig > 0.f ? a.sigv[s] / totsig : 1.f;
    }
    const float skyadd = a.skyadd[s];
    for (int q = tid; q < M2; q += BT) {
      float v = img[q];
      if (totsig > 0.f) v *= scale;
      if (skyadd > 0.f) v += skyadd;
      if (a.skyarr) v += a.skyarr[(long)s * M2 + q];
      img[q] = v;
    }
    __syncthreads();
    if (a.opticalBinning == 1) {   // opticallyBin (:819-850), serial
      if (tid == 0) {
        const int ncen = a.ncen;
        const int st = (M - ncen) / 2, en = M - st;
        float v0 = 0.f;
        for (int i = st; i < en; ++i) v0 += img[M * st + i];
        if (st != 0)
          for (int i = 0; i < ncen; ++i) img[i] = 0.f;
        for (int i = st; i < en; ++i)
          for (int j = (st == 0 ? 1 : st); j < en; ++j) img[i - st] += img[j * M + i];
        for (int i = st + 1; i < en; ++i) {
          if (st != 0) img[i - st + ncen] = 0.f;
          for (int j = (st == 0 ? 1 : st); j < en; ++j) img[i - st + ncen] += img[i * M + j];
        }
        img[ncen] = v0;
        for (int i = 2 * ncen; i < M2; ++i) img[i] = 0.f;
      }
      __syncthreads();
      nexposed = 2 * a.ncen;
    }
    const long noff = (long)frame * a.noise_stride + (long)s * M2;
    if (a.poisson_on && a.poisson_on[s]) {
      for (int q = tid; q < nexposed; q += BT) img[q] = poisson_from_normal(img[q], a.zpois[noff + q]);
      __syncthreads();
    }
    // readCCD (:785-817)
    const bool rn = a.readnoise != 0.f && a.zread;
    float thr = a.floor;
    if (a.threshType == 2) {
      float mx = -INFINITY;
      for (int q = tid; q < nexposed; q += BT) {
        float v = img[q] + (rn ? fmaf(a.readnoise, a.zread[noff + q], a.readbg) : a.readbg);
        mx = fmaxf(mx, v);
      }
      mx = block_max(mx, L);
      if (a.floor > 0.f && a.floor <= 1.f) thr = a.floor * fmaxf(mx, 0.f);
    }
    if (a.threshType >= 0 && a.threshType <= 2) {   // other types: the reference leaves the data
      for (int q = tid; q < nexposed; q += BT) {
        float v = img[q] + (rn ? fmaf(a.readnoise, a.zread[noff + q], a.readbg) : a.readbg);
        if (a.threshType == 1) v = v < thr ? 0.f : v;
        else v = v < thr ? 0.f : v - thr;
        img[q] = v;
      }
      __syncthreads();
    }
  } else {
    // ---------------- image mode (centroidsFromImage, :2000-2031) ----------------
    if (a.inputImageDims == 2) {
      int r0, r1, c0, c1;
      if (a.subapLocation == nullptr) {
        r0 = (s / a.nsubx) * M;
        r1 = r0 + M;
        c0 = (s % a.nsubx) * M;
        c1 = c0 + M;
      } else {
        r0 = a.subapLocation[s * 4];
        r1 = a.subapLocation[s * 4 + 1];
        c0 = a.subapLocation[s * 4 + 2];
        c1 = a.subapLocation[s * 4 + 3];
      }
      const int hgt = r1 - r0, wid = c1 - c0;
      const float *prev = a.bimg + (long)frame * a.bimg_stride + (long)s * M2;
      for (int q = tid; q < M2; q += BT) img[q] = prev[q];
      __syncthreads();
      for (int q = tid; q < hgt * wid; q += BT) {
        int j = q / wid, k = q % wid;
        img[j * M + k] = a.inputImage[(long)(r0 + j) * a.inputImagenpxlx + c0 + k];
      }
      nexposed = hgt * wid;
    } else {
      for (int q = tid; q < M2; q += BT) img[q] = a.inputImage[(long)s * M2 + q];
    }
    __syncthreads();
    if (a.calsource == 0) {
      float thr = a.floor;
      if (a.threshType == 2 && a.floor > 0.f && a.floor <= 1.f) {   // relative floor: f * max
        float mx = -INFINITY;
        for (int q = tid; q < nexposed; q += BT) mx = fmaxf(mx, img[q]);
        thr = a.floor * fmaxf(block_max(mx, L), 0.f);
      }
      for (int q = tid; q < nexposed; q += BT) {
        float v = img[q];
        if (a.threshType == 0 || a.threshType == 2) v = v < thr ? 0.f : v - thr;
        else if (a.threshType == 1) v = v < thr ? 0.f : v;
        img[q] = v;
      }
      __syncthreads();
    }
  }

  // ---------------- brightest pixels (:1544-1581), power (:1893-1897) ----------------
  if (a.calsource == 0) {
    int ub = a.useBrightestArr ? a.useBrightestArr[s] : a.useBrightest;
    bool subtract = false;
    if (ub < 0) {
      ub = -ub;
      subtract = true;
    }
    if (ub > 0 && ub < nexposed) {
      int n2 = 1;
      while (n2 < nexposed) n2 <<= 1;
      float *srt = reinterpret_cast<float *>(L.c);
      for (int q = tid; q < n2; q += BT) srt[q] = q < nexposed ? img[q] : INFINITY;
      __syncthreads();
      lds_sort(srt, n2);
      const float thr = srt[nexposed - ub];
      float sub = 0.f;
      if (subtract) {
        int k = nexposed - ub - 1;
        while (k >= 0 && srt[k] == thr) --k;
        sub = k >= 0 ? srt[k] : 0.f;
      }
      __syncthreads();
      for (int q = tid; q < nexposed; q += BT) {
        float v = img[q];
        img[q] = v < thr ? 0.f : (subtract ? v - sub : v);
      }
      __syncthreads();
    }
  }
  if (a.pxlPower == 2.f) {
    for (int q = tid; q < nexposed; q += BT) img[q] = img[q] * img[q];
    __syncthreads();
  } else if (a.pxlPower != 1.f && a.pxlPower != 0.f) {
    for (int q = tid; q < nexposed; q += BT) img[q] = powf(img[q], a.pxlPower);
    __syncthreads();
  }

  float *bout = a.bimg + (long)frame * a.bimg_stride + (long)s * M2;
  if (a.imageOnly) {
    for (int q = tid; q < M2; q += BT) bout[q] = img[q];
    return;
  }

  // ---------------- correlation (:1899-1926) ----------------
  const int C = a.corrsize;
  const float *cimg = img;   // image the centroid is computed from
  float *cbuf = nullptr;     // the correlation image (written to corrimg)
  if (a.corrMode == 1 && a.corrK) {
    const int o = (C - M) / 2;
    for (int q = tid; q < C * C; q += BT) {
      int i = q / C - o, j = q % C - o;
      L.c[q] = make_float2((i >= 0 && j >= 0 && i < M && j < M) ? img[i * M + j] : 0.f, 0.f);
    }
    __syncthreads();
    lds_fft2(L.c, C, -1, L);
    const float2 *K = a.corrK + (long)s * C * C;
    for (int q = tid; q < C * C; q += BT) {
      float2 u = L.c[q], k = K[q];
      L.c[q] = make_float2(fmaf(u.x, k.x, -u.y * k.y), fmaf(u.x, k.y, u.y * k.x));
    }
    __syncthreads();
    lds_fft2(L.c, C, +1, L);
    for (int q = tid; q < C * C; q += BT) L.h[q] = L.c[q].x;
    __syncthreads();
    cbuf = L.h;
  } else if (a.corrMode > 1 && a.corrMode < 6 && a.corrPattern) {
    const float *pat = a.corrPattern + (long)s * C * C;
    const int ncen = a.ncen;
    for (int q = tid; q < C * C; q += BT) L.h[q] = 0.f;
    __syncthreads();
    const int miny = C < M ? C : M;
    const int half = (M - C) / 2;   // C division truncates toward zero, as in the reference
    for (int q = tid; q < ncen * ncen; q += BT) {
      const int i = q / ncen, j = q % ncen;
      if (a.corrMode == 2 || a.corrMode == 4 || a.corrMode == 5) {
        // mode 1/3 of diffCorrelation (corrMode 2/4) and bruteCorrelation (corrMode 5)
        int offy = i - ncen / 2, my = miny;
        int srefy = -offy - half;
        if (srefy < 0) srefy = 0;
        int siny = half + offy;
        if (siny < 0) siny = 0;
        if (siny + my > M) my = M - siny;
        if (srefy + my > C) my = C - srefy;
        int offx = j - ncen / 2, mx = miny;
        int srefx = -offx - half;
        if (srefx < 0) srefx = 0;
        int sinx = half + offx;
        if (sinx < 0) sinx = 0;
        if (sinx + mx > M) mx = M - sinx;
        if (srefx + mx > C) mx = C - srefx;
        float sacc = 0.f, tot = 0.f, tot2 = 0.f;
        for (int y = 0; y < my; ++y)
          for (int x = 0; x < mx; ++x) {
            float cc = pat[(srefy + y) * C + srefx + x];
            float bb = img[(siny + y) * M + sinx + x];
            if (a.corrMode == 5) {
              sacc += cc * bb;
            } else {
              float d = cc - bb;
              sacc += d * d;
              tot += bb;
              tot2 += cc;
            }
          }
        const int oi = i + (C - ncen) / 2, oj = j + (C - ncen) / 2;
        if (a.corrMode == 2) {
          if (mx != 0 && my != 0) L.h[oi * C + oj] = sacc / (mx * my);
        } else if (a.corrMode == 4) {
          if (tot != 0 && tot2 != 0) L.h[oi * C + oj] = sacc / (tot * tot2);
        } else {
          if (mx != 0 && my != 0) L.h[oi * C + oj] = sacc;
        }
      } else {
        // diffCorrelation's fallback loop (corrMode 3, reference mode 2, :1309-1336)
        const int my = C - abs(ncen / 2 - i);
        const int mx = C - abs(ncen / 2 - j);
        float sacc = 0.f;
        for (int y = 0; y < my; ++y)
          for (int x = 0; x < mx; ++x) {
            float d;
            if (i < ncen / 2) {
              if (j < ncen / 2) d = pat[y * C + x] - img[(ncen / 2 - i + y) * M + ncen / 2 - j + x];
              else d = pat[y * C + j - ncen / 2 + x] - img[(ncen / 2 - i + y) * M + x];
            } else {
              if (j < ncen / 2) d = pat[(i - ncen / 2 + y) * C + x] - img[y * M + ncen / 2 - j + x];
              else d = pat[(i - ncen / 2 + y) * C + j - ncen / 2 + x] - img[y * M + x];
            }
            sacc += fabsf(d);
          }
        L.h[(i + (M - ncen) / 2) * C + j + (M - ncen) / 2] = sacc * sacc / (my * mx);
      }
    }
    __syncthreads();
    cbuf = L.h;
  }
  if (cbuf) {   // thresholdCorrelation (:1488-1505)
    float t = a.corrThresh;
    if (t <= 1.f && t > 0.f) {
      float mx = 0.f;
      for (int q = tid; q < C * C; q += BT) mx = fmaxf(mx, cbuf[q]);
      mx = block_max(mx, L);
      t = mx * t;
    }
    for (int q = tid; q < C * C; q += BT) cbuf[q] = cbuf[q] < t ? 0.f : cbuf[q] - t;
    __syncthreads();
    cimg = cbuf;
  }

  // ---------------- centroid ----------------
  const int N = (cbuf ? C : a.corrsize);
  float cx = 0.f, cy = 0.f;
  if (a.opticalBinning == 1) {
    if (tid == 0) {
      const int ncen = a.ncen;
      float sx = 0, sy = 0, ax = 0, ay = 0;
      for (int i = 0; i < ncen; ++i) {
        sx += img[i];
        ax += img[i] * i;
        sy += img[i + ncen];
        ay += img[i + ncen] * i;
      }
      // computeOBCents (:852-872): with no positive flux the weighted sum itself is returned
      cx = sx > 0 ? ax / sx - (ncen - 1) / 2.f : ax;
      cy = sy > 0 ? ay / sy - (ncen - 1) / 2.f : ay;
    }
  } else if (a.doFit) {
    if (tid == 0)
      fit(const_cast<float *>(cimg), N, a.ncen, a.fitsize, a.fitMx, a.gaussianFit, a.gaussianMinVal,
          a.gaussianReplaceVal, a.corrMode, &cx, &cy);
  } else {
    const float *wt = nullptr;
    if (a.centWeightDim == 2) wt = a.centWeight;
    else if (a.centWeightDim == 4 && a.centWeight) wt = a.centWeight + (long)s * N * N;
    int cob = a.corrMode;
    if (cob == 5) cob = 1;
    cog(cimg, N, a.ncen, wt, cob, &cx, &cy, L);
  }
  __syncthreads();
  for (int q = tid; q < M2; q += BT) bout[q] = img[q];
  if (cbuf && a.corrimg) {
    float *co = a.corrimg + (long)frame * a.corr_stride + (long)s * C * C;
    for (int q = tid; q < C * C; q += BT) co[q] = cbuf[q];
  }
  if (tid == 0) {
    if (a.calData) {
      int lo = a.calBounds[s * 2];
      int n = a.calBounds[s * 2 + 1] - lo;
      cx = interp_cal(cx, n, a.calData + (long)s * a.calNSteps + lo, a.calSteps + lo);
      lo = a.calBounds[(a.nsubaps + s) * 2];
      n = a.calBounds[(a.nsubaps + s) * 2 + 1] - lo;
      cy = interp_cal(cy, n, a.calData + (long)(a.nsubaps + s) * a.calNSteps + lo, a.calSteps + lo);
    } else if (a.calCoeff) {
      const float *c0 = a.calCoeff + (long)(s * 2) * a.calNCoeff;
      const float *c1 = a.calCoeff + (long)(s * 2 + 1) * a.calNCoeff;
      float r = 0.f, xc = 1.f;
      for (int i = 0; i < a.calNCoeff; ++i) { r += xc * c0[i]; xc *= cx; }
      cx = r;
      r = 0.f;
      xc = 1.f;
      for (int i = 0; i < a.calNCoeff; ++i) { r += xc * c1[i]; xc *= cy; }
      cy = r;
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

hipError_t launch_backend(const BackArgs &a, hipStream_t st) {
  if (a.nitems <= 0) return hipSuccess;
  const size_t bytes = BackLdsSizes(a).bytes();
  if (bytes > 64 * 1024) {   // the largest geometries (side 64) need more than the default limit
    static bool raised = false;
    if (!raised) {
      const hipError_t e = hipFuncSetAttribute((const void *)backend_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)(160 * 1024));
      if (e != hipSuccess) return e;
      raised = true;
    }
  }
  hipLaunchKernelGGL(backend_kernel, dim3((unsigned)a.nitems), dim3(BT), (unsigned)bytes, st, a);
  return hipGetLastError();
}

__global__ void inactive_kernel(InactiveArgs a) {
  const long total = (long)a.nframes * a.ninactive;
  const int M2 = a.nimg * a.nimg;
  for (long it = blockIdx.x; it < total; it += gridDim.x) {
    const int frame = (int)(it / a.ninactive);
    const int s = a.inactive[it - (long)frame * a.ninactive];
    if (a.imageOnly == 0) {
      if (threadIdx.x < 2) a.cents[(long)frame * a.cents_stride + s * 2 + threadIdx.x] = 0.f;
    } else {
      float *b = a.bimg + (long)frame * a.bimg_stride + (long)s * M2;
      const float *z = a.zread ? a.zread + (long)frame * a.noise_stride + (long)s * M2 : nullptr;
      const bool rn = a.readnoise != 0.f && z;
      float thr = a.floor;
      if (a.threshType == 2 && a.floor > 0.f && a.floor <= 1.f) {   // relative floor: f * max
        __shared__ float red[BT];
        float mx = -INFINITY;
        for (int q = threadIdx.x; q < M2; q += blockDim.x) mx = fmaxf(mx, rn ? fmaf(a.readnoise, z[q], a.readbg) : a.readbg);
        red[threadIdx.x] = mx;
        __syncthreads();
        for (int w = BT / 2; w > 0; w >>= 1) {
          if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
          __syncthreads();
        }
        thr = a.floor * fmaxf(red[0], 0.f);
        __syncthreads();
      }
      for (int q = threadIdx.x; q < M2; q += blockDim.x) {
        float v = 0.f;
        if (a.threshType >= 0 && a.threshType <= 2) {   // other types: the reference leaves the data
          v += rn ? fmaf(a.readnoise, z[q], a.readbg) : a.readbg;
          if (a.threshType == 1) v = v < thr ? 0.f : v;
          else v = v < thr ? 0.f : v - thr;
        }
        b[q] = v;
      }
    }
  }
}

// Zero slopes of inactive subapertures: one thread per (frame, subaperture).
__global__ void inactive_cents_kernel(InactiveArgs a) {
  const long total = (long)a.nframes * a.ninactive;
  const long it = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= total) return;
  const int frame = (int)(it / a.ninactive);
  const int s = a.inactive[it - (long)frame * a.ninactive];
  *reinterpret_cast<float2 *>(a.cents + (long)frame * a.cents_stride + s * 2) = make_float2(0.f, 0.f);
}

hipError_t launch_inactive(const InactiveArgs &a, hipStream_t st) {
  const long total = (long)a.nframes * a.ninactive;
  if (total <= 0) return hipSuccess;
  if (a.imageOnly == 0) {
    hipLaunchKernelGGL(inactive_cents_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
  }
  long blocks = total < 65535 ? total : 65535;
  hipLaunchKernelGGL(inactive_kernel, dim3((unsigned)blocks), dim3(256), 0, st, a);
  return hipGetLastError();
}

// Zero slopes of the inactive subapertures of several WFS in one launch (multi-WFS step).
__global__ void batch_inactive_kernel(const InactiveDesc *d, int nd, long total) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  int lo = 0, hi = nd - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].start <= t) lo = mid;
    else hi = mid - 1;
  }
  const InactiveDesc q = d[lo];
  const long it = t - q.start;
  const int frame = (int)(it / q.ninactive);
  const int s = q.inactive[it - (long)frame * q.ninactive];
  *reinterpret_cast<float2 *>(q.cents + (long)frame * q.cents_stride + s * 2) = make_float2(0.f, 0.f);
}

hipError_t launch_batch_inactive(const InactiveDesc *d, int nd, long total, hipStream_t st) {
  if (total <= 0 || nd <= 0) return hipSuccess;
  hipLaunchKernelGGL(batch_inactive_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, d, nd, total);
  return hipGetLastError();
}

}  // namespace dasp
