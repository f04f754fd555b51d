// atmos.hip - the phase producer upstream of the WFS engine on the device: createPupilPhs for a
// natural guide star (the layers' bilinear windows summed in float64, piston, pupil, wavelength
// scale), makeLayer as a ring-buffer update, and the deformable mirror (actuator coupling, the
// natural-cubic-spline surface as two products with a fixed operator, the line-of-sight add with
// pupil and piston). C ABI: include/dasp_atmos.h.
//
// All of it is HBM-bound elementwise and reduction work (a 1280 x 1280 pupil is 13 MB of float64
// phase per layer window read), so the kernels are plain coalesced row passes: one block of 256
// threads per 256 pixels of one row, the row index in blockIdx.y (no division per pixel).
// Rounding follows the reference: every product and sum of linearshift and of the line-of-sight
// arithmetic is rounded separately (no fused multiply-add), as the reference's C and NumPy code
// round them.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/dasp_atmos.h"
#include "../../include/dasp_cent.h"

namespace {

constexpr int kRowBlock = 256;

thread_local std::string a_err;
int afail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  a_err = buf;
  return code;
}

#define ACHK(expr)                                                                              \
  do {                                                                                          \
    hipError_t _e = (expr);                                                                     \
    if (_e != hipSuccess) return afail(DASP_CENT_EHIP, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

struct LayerArgs {
  dasp_atmos_layer l[DASP_ATMOS_MAX_LAYERS];
  int nl;
};

// block sum of one double per thread (256 threads), result in thread 0
__device__ double block_sum(double v) {
  __shared__ double red[kRowBlock / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kRowBlock / 64; ++i) t += red[i];
  return t;
}

// ring-buffer index: logical r + origin o, wrapped once (r < n and o < n)
__device__ __forceinline__ int wrap(int r, int n) { return r >= n ? r - n : r; }

// Pass 1: the layers' windows summed (linearshift, cmod/src/interpmodule.c:1293-1300, products and
// sums rounded one by one), the sum written to sum64, and per-block partial sums of sum * pfn
// for the piston.
__global__ __launch_bounds__(kRowBlock) void atmos_sum_kernel(LayerArgs L, int npup, const double *intrinsic,
                                                              const float *pupfn, double *sum64, double *partial) {
  const int i = blockIdx.y;
  const int j = blockIdx.x * kRowBlock + threadIdx.x;
  double pv = 0.0;
  if (j < npup) {
    const long o = (long)i * npup + j;
    double acc = intrinsic ? intrinsic[o] : 0.0;
    for (int k = 0; k < L.nl; ++k) {
      const dasp_atmos_layer &y = L.l[k];
      const int r = wrap(y.r0 + i + y.oy, y.rows), r1 = wrap(r + 1, y.rows);
      const int c = wrap(y.c0 + j + y.ox, y.cols), c1 = wrap(c + 1, y.cols);
      const double *s = y.screen;
      double v = __dmul_rn(s[(long)r * y.cols + c], (double)y.s1);
      v = __dadd_rn(v, __dmul_rn(s[(long)r1 * y.cols + c], (double)y.s2));
      v = __dadd_rn(v, __dmul_rn(s[(long)r * y.cols + c1], (double)y.s3));
      v = __dadd_rn(v, __dmul_rn(s[(long)r1 * y.cols + c1], (double)y.s4));
      acc = __dadd_rn(acc, v);
    }
    sum64[o] = acc;
    pv = pupfn ? __dmul_rn(acc, (double)pupfn[o]) : acc;
  }
  const double t = block_sum(pv);
  if (threadIdx.x == 0) partial[(long)blockIdx.y * gridDim.x + blockIdx.x] = t;
}

// The partial sums reduced by one block into out[0] = sum / area.
__global__ __launch_bounds__(kRowBlock) void mean_kernel(const double *partial, long n, double area, double *out) {
  double v = 0.0;
  for (long k = threadIdx.x; k < n; k += kRowBlock) v += partial[k];
  const double t = block_sum(v);
  if (threadIdx.x == 0) out[0] = t / area;
}

// Pass 2: piston, pupil, wavelength scale (util/atmos.py:821-832, in that order).
__global__ __launch_bounds__(kRowBlock) void atmos_finish_kernel(int npup, const double *sum64, const double *mean,
                                                                 int remove_piston, const float *pupfn, double scale,
                                                                 double *out64, float *out32) {
  const int i = blockIdx.y;
  const int j = blockIdx.x * kRowBlock + threadIdx.x;
  if (j >= npup) return;
  const long o = (long)i * npup + j;
  double v = sum64[o];
  if (remove_piston) v = __dsub_rn(v, mean[0]);
  if (pupfn) v = __dmul_rn(v, (double)pupfn[o]);
  if (scale != 1.0) v = __dmul_rn(v, scale);
  if (out64) out64[o] = v;
  if (out32) out32[o] = (float)v;
}

// makeLayer as a ring update in one launch: the new columns (colInput layout: ncols x rows) at
// logical columns c_first .. c_first + ncols - 1 of every row and the new rows at logical rows
// r_first .. r_first + nrows - 1 of every column. The reference writes the columns first and the
// rows over them, so a column element inside the new rows is left to the row write.
__global__ void ring_update_kernel(double *scr, int rows, int cols, int oy, int ox, const double *csrc, int nc,
                                   int c_first, const double *rsrc, int nr, int r_first) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long ncol_el = (long)rows * nc;
  if (t < ncol_el) {
    const int k = (int)(t / rows), r = (int)(t - (long)k * rows);   // source row k of the new columns
    if (r >= r_first && r < r_first + nr) return;                    // the rows own it
    scr[(long)wrap(r + oy, rows) * cols + wrap(c_first + k + ox, cols)] = csrc[(long)k * rows + r];
    return;
  }
  const long u = t - ncol_el;
  if (u >= (long)cols * nr) return;
  const int k = (int)(u / cols), c = (int)(u - (long)k * cols);
  scr[(long)wrap(r_first + k + oy, rows) * cols + wrap(c + ox, cols)] = rsrc[(long)k * cols + c];
}

// MirrorSurface.fudge (util/dm.py:1723-1740) in float32: the neighbour sum in NumPy's order
// (row above, row below, column left, column right), divided by the neighbour count (4, 3 on the
// edges, 2 at the corners), then a + (d - a) * coupling, each operation rounded.
__global__ void fudge_kernel(const float *a, int n, float coupling, float *out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * n) return;
  const int i = t / n, j = t - (t / n) * n;
  float d = i > 0 ? a[t - n] : 0.f;
  if (i < n - 1) d = __fadd_rn(d, a[t + n]);
  if (j > 0) d = __fadd_rn(d, a[t - 1]);
  if (j < n - 1) d = __fadd_rn(d, a[t + 1]);
  const int cnt = (i > 0) + (i < n - 1) + (j > 0) + (j < n - 1);
  if (n > 1) d = __fdiv_rn(d, (float)cnt);
  out[t] = __fadd_rn(a[t], __fmul_rn(__fsub_rn(d, a[t]), coupling));
}

// Pass 1 of the spline (gslCubSplineIntrp's first dimension): tmp[i][b] = sum_a op[i][a] F[a][b]
// for every pixel row i and actuator column b, in double.
__global__ void spline_rows_kernel(const double *op, const float *F, int npup, int nact, double *tmp) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)npup * nact) return;
  const int i = (int)(t / nact), b = (int)(t - (long)i * nact);
  const double *w = op + (long)i * nact;
  double v = 0.0;
  for (int a = 0; a < nact; ++a) v = fma(w[a], (double)F[(long)a * nact + b], v);
  tmp[t] = v;
}

// Pass 2: surface[i][k] = sum_b tmp[i][b] opT[b][k], cast to float32. A thread takes four rows i
// of one column k: four independent accumulations, each opT element loaded once for the four
// (threads along k read opT rows coalesced; the tmp rows are the same for the whole block).
constexpr int kSplineRows = 4;
__global__ __launch_bounds__(kRowBlock) void spline_cols_kernel(const double *tmp, const double *opT, int npup, int nact,
                                                                float *surface) {
  const int i0 = blockIdx.y * kSplineRows;
  const int k = blockIdx.x * kRowBlock + threadIdx.x;
  if (k >= npup) return;
  const int nr = min(kSplineRows, npup - i0);
  double v[kSplineRows] = {0.0, 0.0, 0.0, 0.0};
  const double *row = tmp + (long)i0 * nact;
  for (int b = 0; b < nact; ++b) {
    const double o = opT[(long)b * npup + k];
#pragma unroll
    for (int r = 0; r < kSplineRows; ++r)
      if (r < nr) v[r] = fma(row[(long)r * nact + b], o, v[r]);
  }
#pragma unroll
  for (int r = 0; r < kSplineRows; ++r)
    if (r < nr) surface[(long)(i0 + r) * npup + k] = (float)v[r];
}

// Line of sight, pass 1: out = base + factor * window (float32 as xinterp_dm's outputData), times
// the pupil, partial sums for the piston.
__global__ __launch_bounds__(kRowBlock) void los_add_kernel(const float *surf, int dmpup, int yoff, int xoff, int npup,
                                                            float factor, int adjust, const double *base64,
                                                            const float *base32, const float *pupfn, float *out,
                                                            double *partial) {
  const int i = blockIdx.y;
  const int j = blockIdx.x * kRowBlock + threadIdx.x;
  double pv = 0.0;
  if (j < npup) {
    const long o = (long)i * npup + j;
    float sel = surf[(long)(i + yoff) * dmpup + (j + xoff)];
    if (adjust) sel = __fmul_rn(sel, factor);
    else if (factor < 0.f) sel = -sel;
    float v = base64 ? (float)base64[o] : (base32 ? base32[o] : 0.f);
    v = (base64 || base32) ? __fadd_rn(v, sel) : sel;
    if (pupfn) v = __fmul_rn(v, pupfn[o]);
    out[o] = v;
    pv = (double)v;
  }
  const double t = block_sum(pv);
  if (threadIdx.x == 0) partial[(long)blockIdx.y * gridDim.x + blockIdx.x] = t;
}

__global__ __launch_bounds__(kRowBlock) void los_piston_kernel(int npup, const double *mean, const float *pupfn,
                                                               float *out) {
  const int i = blockIdx.y;
  const int j = blockIdx.x * kRowBlock + threadIdx.x;
  if (j >= npup) return;
  const long o = (long)i * npup + j;
  out[o] = __fmul_rn(__fsub_rn(out[o], (float)mean[0]), pupfn[o]);
}

int row_blocks(int npup) { return (npup + kRowBlock - 1) / kRowBlock; }

}  // namespace

extern "C" {

const char *dasp_atmos_last_error(void) { return a_err.c_str(); }

long dasp_atmos_work_doubles(int npup) {
  if (npup < 1) return 0;
  return (long)npup * npup + (long)npup * row_blocks(npup) + 1;
}

int dasp_atmos_pupil_phase(const dasp_atmos_layer *layers, int nlayers, int npup, const double *d_intrinsic,
                           const float *d_pupfn, double pupil_area, int remove_piston, double out_scale,
                           double *d_out64, float *d_out32, double *d_work, void *stream) {
  if (npup < 1 || nlayers < 0 || nlayers > DASP_ATMOS_MAX_LAYERS || (nlayers > 0 && !layers))
    return afail(DASP_CENT_EINVAL, "npup >= 1 and 0..%d layers are required", DASP_ATMOS_MAX_LAYERS);
  if (!d_out64 && !d_out32) return afail(DASP_CENT_EINVAL, "d_out64 or d_out32 is required");
  if (!d_work) return afail(DASP_CENT_EINVAL, "d_work (dasp_atmos_work_doubles doubles) is required");
  if (remove_piston && !(pupil_area > 0.0)) return afail(DASP_CENT_EINVAL, "pupil_area must be positive");
  LayerArgs L{};
  L.nl = nlayers;
  for (int k = 0; k < nlayers; ++k) {
    const dasp_atmos_layer &y = layers[k];
    // the window (npup + 1 rows and columns from r0, c0) must lie inside the logical screen and
    // the origin inside the ring, so that one wrap per index suffices
    if (!y.screen || y.rows < npup + 1 || y.cols < npup + 1 || y.r0 < 0 || y.c0 < 0 || y.r0 + npup + 1 > y.rows ||
        y.c0 + npup + 1 > y.cols || y.oy < 0 || y.oy >= y.rows || y.ox < 0 || y.ox >= y.cols)
      return afail(DASP_CENT_EINVAL, "layer %d: window (%d, %d) + %d outside its %d x %d screen, or origin (%d, %d) outside it",
                   k, y.r0, y.c0, npup + 1, y.rows, y.cols, y.oy, y.ox);
    L.l[k] = y;
  }
  hipStream_t st = (hipStream_t)stream;
  const int bx = row_blocks(npup);
  double *sum64 = d_work, *partial = d_work + (long)npup * npup, *mean = partial + (long)npup * bx;
  const dim3 grid(bx, npup);
  hipLaunchKernelGGL(atmos_sum_kernel, grid, dim3(kRowBlock), 0, st, L, npup, d_intrinsic, d_pupfn, sum64, partial);
  ACHK(hipGetLastError());
  if (remove_piston) {
    hipLaunchKernelGGL(mean_kernel, dim3(1), dim3(kRowBlock), 0, st, partial, (long)npup * bx, pupil_area, mean);
    ACHK(hipGetLastError());
  }
  hipLaunchKernelGGL(atmos_finish_kernel, grid, dim3(kRowBlock), 0, st, npup, sum64, mean, remove_piston, d_pupfn,
                     out_scale, d_out64, d_out32);
  ACHK(hipGetLastError());
  return 0;
}

int dasp_atmos_layer_update(double *d_screen, int rows, int cols, int *oy, int *ox, const double *d_cols, int ncols,
                            const double *d_rows, int nrows, int col_add, int row_add, void *stream) {
  if (!d_screen || rows < 1 || cols < 1 || !oy || !ox || ncols < 0 || nrows < 0 || ncols > cols || nrows > rows ||
      (ncols > 0 && !d_cols) || (nrows > 0 && !d_rows) || *oy < 0 || *oy >= rows || *ox < 0 || *ox >= cols)
    return afail(DASP_CENT_EINVAL, "bad screen, origin or new-data arguments");
  hipStream_t st = (hipStream_t)stream;
  // the shift: the logical screen moves so that the old logical (r + nrows, c + ncols) becomes
  // (r, c) when data enters at the end, (r - nrows, c - ncols) when it enters at the start
  int y = *oy, x = *ox;
  y = row_add < 0 ? (y + nrows) % rows : (y - nrows % rows + rows) % rows;
  x = col_add < 0 ? (x + ncols) % cols : (x - ncols % cols + cols) % cols;
  const long n = (long)rows * ncols + (long)cols * nrows;
  if (n > 0) {
    hipLaunchKernelGGL(ring_update_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d_screen, rows, cols, y,
                       x, d_cols, ncols, col_add < 0 ? cols - ncols : 0, d_rows, nrows, row_add < 0 ? rows - nrows : 0);
    ACHK(hipGetLastError());
  }
  *oy = y;
  *ox = x;
  return 0;
}

int dasp_dm_spline_operator(int npup, int nact, double actoffset, double *op) {
  if (npup < 1 || nact < 2 || !op) return afail(DASP_CENT_EINVAL, "npup >= 1, nact >= 2 and op are required");
  const double spacing = (npup - 1.) / (nact - 1 + actoffset * 2.);
  std::vector<double> x2(nact);
  for (int a = 0; a < nact; ++a) x2[a] = a * spacing + spacing * actoffset;
  // the span must cover every pixel; a last knot short of npup - 1 by rounding (a few ulp of
  // (nact - 1) * spacing) is accepted, its cubic taking that last fraction
  const double tol = 1e-9 * npup;
  if (x2[0] > tol || x2[nact - 1] < npup - 1.0 - tol)
    return afail(DASP_CENT_EUNSUPPORTED, "pixels outside the actuator span (actoffset %g): the spline is not defined there",
                 actoffset);
  // natural spline of each unit vector: c from the tridiagonal system (c[0] = c[n-1] = 0), then
  // the cubic on each interval (GSL's cspline formulas)
  const int n = nact, m = n - 2;
  std::vector<double> h(n - 1), c(n), y(n), diag(m > 0 ? m : 1), r(m > 0 ? m : 1);
  for (int k = 0; k < n - 1; ++k) h[k] = x2[k + 1] - x2[k];
  for (int a = 0; a < n; ++a) {
    std::fill(y.begin(), y.end(), 0.0);
    y[a] = 1.0;
    std::fill(c.begin(), c.end(), 0.0);
    if (m > 0) {
      for (int k = 0; k < m; ++k) {
        diag[k] = 2.0 * (h[k] + h[k + 1]);
        r[k] = 3.0 * ((y[k + 2] - y[k + 1]) / h[k + 1] - (y[k + 1] - y[k]) / h[k]);
      }
      for (int k = 1; k < m; ++k) {
        const double w = h[k] / diag[k - 1];
        diag[k] -= w * h[k];
        r[k] -= w * r[k - 1];
      }
      c[m] = r[m - 1] / diag[m - 1];
      for (int k = m - 2; k >= 0; --k) c[k + 1] = (r[k] - h[k + 1] * c[k + 2]) / diag[k];
    }
    int iv = 0;
    for (int k = 0; k < npup; ++k) {
      const double xe = (double)k;
      while (iv < n - 2 && xe >= x2[iv + 1]) ++iv;
      const double hh = h[iv], dx = xe - x2[iv];
      const double b = (y[iv + 1] - y[iv]) / hh - hh * (c[iv + 1] + 2.0 * c[iv]) / 3.0;
      const double d = (c[iv + 1] - c[iv]) / (3.0 * hh);
      op[(long)k * nact + a] = y[iv] + dx * (b + dx * (c[iv] + dx * d));
    }
  }
  return 0;
}

int dasp_dm_surface(const float *d_act, int nact, float coupling, const double *d_op, const double *d_opT, int npup,
                    float *d_surface, void *d_work, void *stream) {
  if (!d_act || nact < 2 || !d_op || !d_opT || npup < 1 || !d_surface || !d_work)
    return afail(DASP_CENT_EINVAL, "d_act, nact >= 2, the operators, d_surface and d_work are required");
  hipStream_t st = (hipStream_t)stream;
  // work: the coupled map (nact^2 floats, rounded up to 8 bytes) then tmp (npup x nact doubles)
  float *F = (float *)d_work;
  double *tmp = (double *)((char *)d_work + (((long)nact * nact * 4 + 7) / 8) * 8);
  const float *src = d_act;
  if (coupling != 0.f) {
    hipLaunchKernelGGL(fudge_kernel, dim3((nact * nact + 255) / 256), dim3(256), 0, st, d_act, nact, coupling, F);
    ACHK(hipGetLastError());
    src = F;
  }
  const long n1 = (long)npup * nact;
  hipLaunchKernelGGL(spline_rows_kernel, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, st, d_op, src, npup, nact,
                     tmp);
  ACHK(hipGetLastError());
  hipLaunchKernelGGL(spline_cols_kernel, dim3(row_blocks(npup), (npup + kSplineRows - 1) / kSplineRows), dim3(kRowBlock),
                     0, st, tmp, d_opT, npup, nact, d_surface);
  ACHK(hipGetLastError());
  return 0;
}

int dasp_dm_line_of_sight(const float *d_surface, int dmpup, int yoff, int xoff, int npup, float factor, int adjust,
                          const double *d_base64, const float *d_base32, const float *d_pupfn, double pupil_sum,
                          int remove_piston, float *d_out32, double *d_work, void *stream) {
  if (!d_surface || !d_out32 || npup < 1 || dmpup < 1 || yoff < 0 || xoff < 0 || yoff + npup > dmpup ||
      xoff + npup > dmpup || !d_work)
    return afail(DASP_CENT_EINVAL, "the %d x %d window at (%d, %d) must lie inside the %d x %d mirror; d_work required",
                 npup, npup, yoff, xoff, dmpup, dmpup);
  if (remove_piston && (!d_pupfn || !(pupil_sum > 0.0)))
    return afail(DASP_CENT_EINVAL, "the piston is removed inside a pupil: d_pupfn and a positive pupil_sum are required");
  hipStream_t st = (hipStream_t)stream;
  const int bx = row_blocks(npup);
  double *partial = d_work + (long)npup * npup, *mean = partial + (long)npup * bx;
  const dim3 grid(bx, npup);
  hipLaunchKernelGGL(los_add_kernel, grid, dim3(kRowBlock), 0, st, d_surface, dmpup, yoff, xoff, npup, factor, adjust,
                     d_base64, d_base32, d_pupfn, d_out32, partial);
  ACHK(hipGetLastError());
  if (remove_piston) {
    hipLaunchKernelGGL(mean_kernel, dim3(1), dim3(kRowBlock), 0, st, partial, (long)npup * bx, pupil_sum, mean);
    ACHK(hipGetLastError());
    hipLaunchKernelGGL(los_piston_kernel, grid, dim3(kRowBlock), 0, st, npup, mean, d_pupfn, d_out32);
    ACHK(hipGetLastError());
  }
  return 0;
}

}  // extern "C"
