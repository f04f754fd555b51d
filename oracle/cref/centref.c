/*
 * centref.c - FP32 C restatement of DASP's per-subaperture centroiding engine
 * (cmod/src/centmodule.c), for testing and as the timed CPU baseline.
 *
 * TEST INFRASTRUCTURE: only tests/ and bench.py's cpu_baseline leg load this library.
 *
 * It follows centroidsFromPhase (cmod/src/centmodule.c:1760-1984) for the options the benchmark
 * workloads use: computeHll (:473-509) with a radix-2 complex FFT of its own in place of FFTW
 * (FFTW3f is not available here), addPowerSpecCentral (:430-471), the LGS spot convolution
 * (:524-564, complex transforms instead of r2c/c2r), binImage with the prepareBinImage weights
 * (:565-651), the theoretical-total normalisation (:1829-1863, :2403-2427), sky, Poisson and read
 * noise from pre-drawn N(0,1) arrays (the same transform as the GPU engine), readCCD floors
 * (:785-817, plus the engine's relative floor option), pixel power, the FFT correlation with its
 * half-complex product (:1403-1505), computeCoG (:873-971) and computeFit (:1069-1188).
 * Work is split over a pthread pool with the reference's static contiguous partition
 * (:2533-2550).
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int nsubx, n, fftsize, clipsize, nimg, ncen;
  const float *pupfn;      /* (S, n, n) */
  const int *subflag;      /* (S) */
  const float *fracSubArea;
  float sig;
  const float *sigArr;     /* NULL or (S) */
  float skybrightness;
  int addPoisson, calsource;
  float readnoise, readbg, noiseFloor;
  int threshType;
  int relativeFloor;       /* engine option on top of threshType 0: floor = noiseFloor * max */
  float pxlPower;
  const float *spotpsf;    /* NULL or (S, P, P) */
  int psfsize;
  int correlation;         /* 0 or 1 (FFT correlation) */
  const float *corrPattern;/* (S, C, C) half-complex */
  int corrsize;
  float corrThresh;
  int parabolicFit;        /* 0 or fit size */
  const float *refCents;   /* NULL or (S, 2) */
} cref_params;

typedef struct {
  float re, im;
} cpx;

/* ---------------- FFT (radix-2, in place, power-of-two n, stride) ---------------- */
static void fft1(cpx *x, int n, int stride, int sign) {
  int i, j = 0, len, k;
  for (i = 1; i < n; ++i) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      cpx t = x[i * stride];
      x[i * stride] = x[j * stride];
      x[j * stride] = t;
    }
  }
  for (len = 2; len <= n; len <<= 1) {
    double ang = sign * 2.0 * M_PI / len;
    for (k = 0; k < len / 2; ++k) {
      float wr = (float)cos(ang * k), wi = (float)sin(ang * k);
      for (i = 0; i < n; i += len) {
        cpx u = x[(i + k) * stride], v = x[(i + k + len / 2) * stride], t;
        t.re = v.re * wr - v.im * wi;
        t.im = v.re * wi + v.im * wr;
        x[(i + k) * stride].re = u.re + t.re;
        x[(i + k) * stride].im = u.im + t.im;
        x[(i + k + len / 2) * stride].re = u.re - t.re;
        x[(i + k + len / 2) * stride].im = u.im - t.im;
      }
    }
  }
}

static void fft2(cpx *x, int n, int sign, int rows_nonzero) {
  int r, c;
  for (r = 0; r < rows_nonzero; ++r) fft1(x + r * n, n, 1, sign);
  for (c = 0; c < n; ++c) fft1(x + c, n, n, sign);
}

/* ---------------- noise (same transform as the engine) ---------------- */
static float poisson_from_normal(float lam, float z) {
  if (!(lam > 0.f)) return 0.f;
  if (lam < 10.f) {
    const float u = 0.5f * erfcf(-z * 0.70710678118654752f);
    float p = expf(-lam), F = p;
    int k = 0;
    while (u > F && k < 40) {
      const float pn = p * lam / (float)(k + 1);
      const float Fn = F + pn;
      if (!(Fn > F)) break;
      ++k;
      p = pn;
      F = Fn;
    }
    return (float)k;
  }
  {
    const float s = sqrtf(lam + 1.0f / 12.0f), s2 = s * s;
    const float g1 = lam / (s2 * s), g2 = (lam - 1.0f / 120.0f) / (s2 * s2), z2 = z * z;
    const float w = z + g1 * (z2 - 1.0f) / 6.0f + g2 * (z2 * z - 3.0f * z) / 24.0f -
                    g1 * g1 * (2.0f * z2 * z - 5.0f * z) / 36.0f;
    const float k = ceilf(lam + s * w - 0.5f);
    return k > 0.f ? k : 0.f;
  }
}

/* ---------------- binning (prepareBinImage + binImage) ---------------- */
static void bin_image(int clip, const float *in, int nimg, float *out, const float *t, const float *nx, const int *idx) {
  int i, j, ii = 0, jj;
  if (clip == nimg) {
    memcpy(out, in, sizeof(float) * clip * clip);
    return;
  }
  memset(out, 0, sizeof(float) * nimg * nimg);
  for (i = 0; i < clip - 1; ++i) {
    jj = 0;
    for (j = 0; j < clip - 1; ++j) {
      out[ii * nimg + jj] += t[i] * t[j] * in[i * clip + j];
      if (jj + 1 < nimg) out[ii * nimg + jj + 1] += t[i] * nx[j] * in[i * clip + j + 1];
      if (ii + 1 < nimg) {
        out[(ii + 1) * nimg + jj] += nx[i] * t[j] * in[(i + 1) * clip + j];
        if (jj + 1 < nimg) out[(ii + 1) * nimg + jj + 1] += nx[i] * nx[j] * in[(i + 1) * clip + j + 1];
      }
      jj += idx[j];
    }
    out[ii * nimg + jj] += t[i] * t[j] * in[i * clip + j];
    if (ii + 1 < nimg) out[(ii + 1) * nimg + jj] += nx[i] * t[j] * in[(i + 1) * clip + j];
    ii += idx[i];
  }
  jj = 0;
  for (j = 0; j < clip - 1; ++j) {
    out[ii * nimg + jj] += t[i] * t[j] * in[i * clip + j];
    if (jj + 1 < nimg) out[ii * nimg + jj + 1] += t[i] * nx[j] * in[i * clip + j + 1];
    jj += idx[j];
  }
  out[ii * nimg + jj] += t[i] * t[j] * in[i * clip + j];
}

/* ---------------- correlation (R2HC x R2HC, HC product, HC2R) ---------------- */
static void r2hc1(float *a, int n, int stride, cpx *tmp) {
  int k;
  for (k = 0; k < n; ++k) {
    tmp[k].re = a[k * stride];
    tmp[k].im = 0.f;
  }
  fft1(tmp, n, 1, -1);
  for (k = 0; k <= n / 2; ++k) a[k * stride] = tmp[k].re;
  for (k = n / 2 + 1; k < n; ++k) a[k * stride] = tmp[n - k].im;
}

static void hc2r1(float *a, int n, int stride, cpx *tmp) {
  int k;
  for (k = 0; k < n; ++k) {
    if (k == 0 || (n % 2 == 0 && k == n / 2)) {
      tmp[k].re = a[k * stride];
      tmp[k].im = 0.f;
    } else if (k < (n + 1) / 2) {
      tmp[k].re = a[k * stride];
      tmp[k].im = a[(n - k) * stride];
    } else {
      tmp[k].re = a[(n - k) * stride];
      tmp[k].im = -a[k * stride];
    }
  }
  fft1(tmp, n, 1, +1);
  for (k = 0; k < n; ++k) a[k * stride] = tmp[k].re;
}

#define B(y, x) pat[(y) * n + (x)]
static void hc_mult(float *a, const float *pat, int n) {
  int i, j, m = n;
  float r1, r2, r3, r4, r5, r6, r7, r8;
  a[0] *= B(0, 0);
  a[n / 2] *= B(0, n / 2);
  a[n * m / 2] *= B(m / 2, 0);
  a[n * m / 2 + n / 2] *= B(m / 2, n / 2);
  for (i = 1; i < (n + 1) / 2; i++) {
    r1 = a[i] * B(0, i) - a[n - i] * B(0, n - i);
    r2 = a[i] * B(0, n - i) + a[n - i] * B(0, i);
    a[i] = r1;
    a[n - i] = r2;
    r3 = a[m / 2 * n + i] * B(m / 2, i) - a[m / 2 * n + n - i] * B(m / 2, n - i);
    r4 = a[m / 2 * n + i] * B(m / 2, n - i) + a[m / 2 * n + n - i] * B(m / 2, i);
    a[m / 2 * n + i] = r3;
    a[m / 2 * n + n - i] = r4;
  }
  for (i = 1; i < (m + 1) / 2; i++) {
    r5 = a[i * n] * B(i, 0) - a[(m - i) * n] * B(m - i, 0);
    r6 = a[i * n] * B(m - i, 0) + a[(m - i) * n] * B(i, 0);
    a[i * n] = r5;
    a[(m - i) * n] = r6;
    r7 = a[i * n + n / 2] * B(i, n / 2) - a[(m - i) * n + n / 2] * B(m - i, n / 2);
    r8 = a[i * n + n / 2] * B(m - i, n / 2) + a[(m - i) * n + n / 2] * B(i, n / 2);
    a[i * n + n / 2] = r7;
    a[(m - i) * n + n / 2] = r8;
    for (j = 1; j < (n + 1) / 2; j++) {
      float aij = a[i * n + j], amn = a[(m - i) * n + n - j], ain = a[i * n + n - j], amj = a[(m - i) * n + j];
      r1 = aij * B(i, j) + amn * B(m - i, n - j) - ain * B(i, n - j) - amj * B(m - i, j);
      r2 = aij * B(m - i, n - j) + amn * B(i, j) + amj * B(i, n - j) + ain * B(m - i, j);
      r3 = aij * B(i, n - j) - amn * B(m - i, j) + ain * B(i, j) - amj * B(m - i, n - j);
      r4 = aij * B(m - i, j) - amn * B(i, n - j) + amj * B(i, j) - ain * B(m - i, n - j);
      a[i * n + j] = r1;
      a[(m - i) * n + n - j] = r2;
      a[i * n + n - j] = r3;
      a[(m - i) * n + j] = r4;
    }
  }
}
#undef B

/* ---------------- centroids ---------------- */
static void cog(int nimg, int ncen, const float *img, float *cx, float *cy) {
  int i, j, s0 = (nimg - ncen) / 2, e0 = nimg - s0;
  float sum = 0.f, x = 0.f, y = 0.f;
  for (i = s0; i < e0; ++i)
    for (j = s0; j < e0; ++j) {
      float t = img[i * nimg + j];
      sum += t;
      x += t * j;
      y += t * i;
    }
  if (sum > 0) {
    *cx = x / sum - (nimg - 1) / 2.f;
    *cy = y / sum - (nimg - 1) / 2.f;
  } else {
    *cx = *cy = 0.f;
  }
}

static void fit_matrix(int fs, float *out) {
  double A[6][12];
  int r, c, x, y, col;
  for (r = 0; r < 6; ++r)
    for (c = 0; c < 12; ++c) A[r][c] = (c - 6 == r) ? 1.0 : 0.0;
  for (y = 0; y < fs; ++y)
    for (x = 0; x < fs; ++x) {
      double v[6] = {(double)x * x, (double)x * y, (double)y * y, (double)x, (double)y, 1.0};
      for (r = 0; r < 6; ++r)
        for (c = 0; c < 6; ++c) A[r][c] += v[r] * v[c];
    }
  for (col = 0; col < 6; ++col) {
    int piv = col;
    double inv;
    for (r = col + 1; r < 6; ++r)
      if (fabs(A[r][col]) > fabs(A[piv][col])) piv = r;
    if (piv != col)
      for (c = 0; c < 12; ++c) {
        double t = A[piv][c];
        A[piv][c] = A[col][c];
        A[col][c] = t;
      }
    inv = 1.0 / A[col][col];
    for (c = 0; c < 12; ++c) A[col][c] *= inv;
    for (r = 0; r < 6; ++r)
      if (r != col) {
        double f = A[r][col];
        for (c = 0; c < 12; ++c) A[r][c] -= f * A[col][c];
      }
  }
  for (r = 0; r < 5; ++r)
    for (c = 0; c < 6; ++c) out[r * 6 + c] = (float)A[r][6 + c];
}

static void fit(int nimg, int ncen, int fitsize, const float *img, const float *fm, float *cx, float *cy) {
  int i, mxpos = 0, nc = (nimg - ncen) / 2, sx, sy, x, y;
  float sum = 0.f, mx = 0.f, vec[6] = {0, 0, 0, 0, 0, 0}, res[5], X, Y;
  if (fitsize < 3) fitsize = 3;
  for (i = 0; i < nimg * nimg; ++i) {
    sum += img[i];
    if (img[i] > mx) {
      mx = img[i];
      mxpos = i;
    }
  }
  if (!(sum >= 0.f)) {
    *cx = *cy = 0.f;
    return;
  }
  sx = mxpos % nimg - fitsize / 2;
  sy = mxpos / nimg - fitsize / 2;
  if (sx < nc) sx = nc;
  if (sy < nc) sy = nc;
  if (sx > nc + ncen - fitsize) sx = nc + ncen - fitsize;
  if (sy > nc + ncen - fitsize) sy = nc + ncen - fitsize;
  for (y = 0; y < fitsize; ++y)
    for (x = 0; x < fitsize; ++x) {
      float v = img[(y + sy) * nimg + x + sx];
      vec[0] += v * x * x;
      vec[1] += v * x * y;
      vec[2] += v * y * y;
      vec[3] += v * x;
      vec[4] += v * y;
      vec[5] += v;
    }
  for (i = 0; i < 5; ++i) {
    int k;
    float t = 0.f;
    for (k = 0; k < 6; ++k) t += fm[i * 6 + k] * vec[k];
    res[i] = t;
  }
  X = (res[1] * res[4] / (2 * res[2]) - res[3]) / (2.f * res[0] - res[1] * res[1] / (2.f * res[2]));
  Y = -(res[4] + res[1] * X) / (2.f * res[2]);
  if (X < 0 || Y < 0 || X >= fitsize || Y >= fitsize) {
    X = (float)(mxpos % nimg - sx);
    Y = (float)(mxpos / nimg - sy);
  }
  Y -= ncen / 2.f - 0.5f;
  X -= ncen / 2.f - 0.5f;
  X += sx + ncen / 2.f - nimg / 2.f;
  Y += sy + ncen / 2.f - nimg / 2.f;
  *cx = X;
  *cy = Y;
}

/* ---------------- driver ---------------- */
typedef struct {
  const cref_params *p;
  const float *phs, *zpois, *zread;
  float *cents, *bimg, *corrimg;
  int s0, s1;
  const float *tilt, *t, *nx, *totsig, *fitmx;
  const int *idx;
  const cpx *otf;
} job_t;

static void *worker(void *arg) {
  job_t *jb = (job_t *)arg;
  const cref_params *p = jb->p;
  const int n = p->n, F = p->fftsize, Q = p->clipsize, M = p->nimg, M2 = M * M;
  const int P = p->spotpsf ? p->psfsize : Q, npup = p->nsubx * n;
  const int C = p->correlation ? p->corrsize : M;
  int fs = p->spotpsf ? (F > P ? F : P) : F;
  cpx *fa = (cpx *)malloc(sizeof(cpx) * fs * fs);
  float *hll = (float *)malloc(sizeof(float) * P * P);
  float *corr = (float *)malloc(sizeof(float) * C * C);
  cpx *tmp = (cpx *)malloc(sizeof(cpx) * (C > P ? C : P));
  int s;
  for (s = jb->s0; s < jb->s1; ++s) {
    const int sy = s / p->nsubx, sx = s % p->nsubx;
    float *img = jb->bimg + (long)s * M2;
    int i, j;
    float tot, sigv;
    if (p->subflag[s] != 1) {
      jb->cents[2 * s] = jb->cents[2 * s + 1] = 0.f;
      continue;
    }
    memset(fa, 0, sizeof(cpx) * F * F);
    for (i = 0; i < n; ++i)
      for (j = 0; j < n; ++j) {
        float th = jb->phs[(long)(sy * n + i) * npup + sx * n + j] - jb->tilt[i * n + j];
        float pu = p->pupfn[((long)s * n + i) * n + j];
        fa[i * F + j].re = pu * (float)cos(th);
        fa[i * F + j].im = pu * (float)sin(th);
      }
    fft2(fa, F, -1, n);
    /* addPowerSpecCentral */
    memset(hll, 0, sizeof(float) * P * P);
    {
      int c, rs = 0, ro = 0;
      if (P > F) {
        rs = (P - F) / 2;
        c = F;
      } else {
        ro = (F - P) / 2;
        c = P;
      }
      for (i = 0; i < c; ++i)
        for (j = 0; j < c; ++j) {
          cpx v = fa[(ro + i) * F + ro + j];
          hll[(c - 1 - i + rs) * P + c - 1 - j + rs] = v.re * v.re + v.im * v.im;
        }
    }
    if (p->spotpsf) {   /* doConvolution with complex transforms */
      const cpx *K = jb->otf + (long)s * P * P;
      for (i = 0; i < P * P; ++i) {
        fa[i].re = hll[i];
        fa[i].im = 0.f;
      }
      fft2(fa, P, -1, P);
      for (i = 0; i < P * P; ++i) {
        cpx u = fa[i];
        fa[i].re = u.re * K[i].re - u.im * K[i].im;
        fa[i].im = u.re * K[i].im + u.im * K[i].re;
      }
      fft2(fa, P, +1, P);
      {
        int off = (P - Q) / 2;
        for (i = 0; i < Q; ++i)
          for (j = 0; j < Q; ++j) hll[i * Q + j] = fa[(i + off) * P + j + off].re;
      }
    }
    bin_image(Q, hll, M, img, jb->t, jb->nx, jb->idx);
    tot = jb->totsig[s];
    sigv = p->sigArr ? p->sigArr[s] : p->sig * p->fracSubArea[s];
    if (tot > 0)
      for (i = 0; i < M2; ++i) img[i] *= sigv / tot;
    {
      float sky = p->skybrightness * p->fracSubArea[s];
      if (sky > 0)
        for (i = 0; i < M2; ++i) img[i] += sky;
      if ((p->addPoisson == 1 && p->calsource == 0) && (tot > 0 || sky > 0) && jb->zpois)
        for (i = 0; i < M2; ++i) img[i] = poisson_from_normal(img[i], jb->zpois[(long)s * M2 + i]);
    }
    {
      float rn = p->calsource == 0 ? p->readnoise : 0.f, thr = p->noiseFloor, mx = -INFINITY;
      if (p->threshType == 0 || p->threshType == 1) {   /* other types: readCCD leaves the data */
        for (i = 0; i < M2; ++i) {
          img[i] += (rn != 0.f && jb->zread) ? rn * jb->zread[(long)s * M2 + i] + p->readbg : p->readbg;
          if (img[i] > mx) mx = img[i];
        }
        if (p->threshType == 0 && p->relativeFloor && thr > 0.f && thr <= 1.f) thr = thr * (mx > 0 ? mx : 0);
        for (i = 0; i < M2; ++i) {
          if (p->threshType == 1) img[i] = img[i] < thr ? 0.f : img[i];
          else img[i] = img[i] < thr ? 0.f : img[i] - thr;
        }
      }
    }
    if (p->pxlPower == 2.f)
      for (i = 0; i < M2; ++i) img[i] *= img[i];
    else if (p->pxlPower != 1.f && p->pxlPower != 0.f)
      for (i = 0; i < M2; ++i) img[i] = powf(img[i], p->pxlPower);
    if (p->correlation && p->corrPattern) {
      const int o = (C - M) / 2;
      memset(corr, 0, sizeof(float) * C * C);
      for (i = 0; i < M; ++i) memcpy(&corr[(o + i) * C + o], &img[i * M], sizeof(float) * M);
      for (i = 0; i < C; ++i) r2hc1(corr + i * C, C, 1, tmp);
      for (j = 0; j < C; ++j) r2hc1(corr + j, C, C, tmp);
      hc_mult(corr, p->corrPattern + (long)s * C * C, C);
      for (i = 0; i < C; ++i) hc2r1(corr + i * C, C, 1, tmp);
      for (j = 0; j < C; ++j) hc2r1(corr + j, C, C, tmp);
      {
        float t = p->corrThresh, mx = 0.f;
        if (t <= 1.f && t > 0.f) {
          for (i = 0; i < C * C; ++i)
            if (corr[i] > mx) mx = corr[i];
          t *= mx;
        }
        for (i = 0; i < C * C; ++i) corr[i] = corr[i] < t ? 0.f : corr[i] - t;
      }
      if (jb->corrimg) memcpy(jb->corrimg + (long)s * C * C, corr, sizeof(float) * C * C);
      if (p->parabolicFit) fit(C, p->ncen, p->parabolicFit, corr, jb->fitmx, &jb->cents[2 * s], &jb->cents[2 * s + 1]);
      else cog(C, p->ncen, corr, &jb->cents[2 * s], &jb->cents[2 * s + 1]);
    } else if (p->parabolicFit) {
      fit(M, p->ncen, p->parabolicFit, img, jb->fitmx, &jb->cents[2 * s], &jb->cents[2 * s + 1]);
    } else {
      cog(M, p->ncen, img, &jb->cents[2 * s], &jb->cents[2 * s + 1]);
    }
    if (p->refCents) {
      jb->cents[2 * s] -= p->refCents[2 * s];
      jb->cents[2 * s + 1] -= p->refCents[2 * s + 1];
    }
  }
  free(fa);
  free(hll);
  free(corr);
  free(tmp);
  return NULL;
}

/* One readout of every subaperture with nthreads threads. zpois/zread: (S, nimg, nimg) N(0,1) or
 * NULL. Returns 0 on success. */
int cref_run(const cref_params *p, const float *phs, const float *zpois, const float *zread, float *cents,
             float *bimg, float *corrimg, int nthreads) {
  const int S = p->nsubx * p->nsubx, n = p->n, F = p->fftsize, Q = p->clipsize, M = p->nimg;
  const int P = p->spotpsf ? p->psfsize : Q;
  float *tilt = (float *)malloc(sizeof(float) * n * n);
  float *t = (float *)malloc(sizeof(float) * Q), *nx = (float *)malloc(sizeof(float) * Q);
  int *idx = (int *)malloc(sizeof(int) * Q);
  float *totsig = (float *)malloc(sizeof(float) * S);
  float fitmx[30];
  cpx *otf = NULL;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
  job_t *jobs = (job_t *)malloc(sizeof(job_t) * nthreads);
  int i, j, cnt = 0, left = S, start = 0;
  const float tmpf = (float)(M_PI / F * (F + 1.));
  for (i = 0; i < n; ++i)
    for (j = 0; j < n; ++j) tilt[i * n + j] = tmpf * (i + j + 1 - n);
  for (i = 0; i < Q; ++i) {   /* prepareBinImage */
    cnt += M;
    if (cnt >= Q) {
      int ex = cnt - Q;
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
  for (i = 0; i < S; ++i) {   /* updateTotSigArr */
    float tt = 0.f;
    for (j = 0; j < n * n; ++j) tt += p->pupfn[(long)i * n * n + j];
    tt *= (float)(F * F);
    if (p->spotpsf) {
      float ts = 0.f;
      for (j = 0; j < P * P; ++j) ts += p->spotpsf[(long)i * P * P + j];
      tt *= ts * P * P;
    }
    totsig[i] = tt;
  }
  if (p->spotpsf) {   /* prepareSpotPsf: OTF of the quadrant-swapped spot */
    otf = (cpx *)malloc(sizeof(cpx) * (long)S * P * P);
    for (i = 0; i < S; ++i) {
      cpx *o = otf + (long)i * P * P;
      const float *sp = p->spotpsf + (long)i * P * P;
      int r, c;
      for (r = 0; r < P; ++r)
        for (c = 0; c < P; ++c) {
          o[r * P + c].re = sp[((r + P / 2) % P) * P + (c + P / 2) % P];
          o[r * P + c].im = 0.f;
        }
      fft2(o, P, -1, P);
    }
  }
  if (p->parabolicFit) fit_matrix(p->parabolicFit < 3 ? 3 : p->parabolicFit, fitmx);
  if (nthreads < 1) nthreads = 1;
  for (i = 0; i < nthreads; ++i) {
    int cntS = left / (nthreads - i);
    jobs[i].p = p;
    jobs[i].phs = phs;
    jobs[i].zpois = zpois;
    jobs[i].zread = zread;
    jobs[i].cents = cents;
    jobs[i].bimg = bimg;
    jobs[i].corrimg = corrimg;
    jobs[i].s0 = start;
    jobs[i].s1 = start + cntS;
    jobs[i].tilt = tilt;
    jobs[i].t = t;
    jobs[i].nx = nx;
    jobs[i].idx = idx;
    jobs[i].totsig = totsig;
    jobs[i].fitmx = fitmx;
    jobs[i].otf = otf;
    start += cntS;
    left -= cntS;
  }
  for (i = 1; i < nthreads; ++i) pthread_create(&th[i], NULL, worker, &jobs[i]);
  worker(&jobs[0]);
  for (i = 1; i < nthreads; ++i) pthread_join(th[i], NULL);
  free(tilt);
  free(t);
  free(nx);
  free(idx);
  free(totsig);
  free(otf);
  free(th);
  free(jobs);
  return 0;
}
