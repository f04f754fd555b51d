"""FP64 NumPy oracle for DASP's Shack-Hartmann image formation and centroiding path.

TEST INFRASTRUCTURE. This module is the checker for the HIP product in ``dasp_amd``. Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import it.
The product never imports it and never falls back to it.

Each function restates one piece of the reference engine ``cmod/src/centmodule.c`` (and the
Python helpers around it) in float64, citing the lines it follows. Citations are relative to the
agb32/dasp source tree.

How each function is pinned. The reference holds no numeric golden vectors for this path, and it
cannot be built here (FFTW3f, GSL, Python 2 C-API). Each function is therefore checked against
mathematics, the reference's C text read as closed forms, or an independent route; the checks
live in tests/test_oracle_pins.py (P) and tests/test_oracle_pins_detector.py (D), where every
pin of D also rejects deliberately planted mistakes (comparisons, dropped terms, off-by-one
windows, shift directions) in a copy of the function:
  compute_hll, tilt_fn,  (P) brute-force 2-D DFT on small inputs; Parseval (sum = F^2 * sum(pup));
  subap_windows          flat phase + symmetric pupil gives a centred spot (centroid 0); a ramp
                         of one FFT pixel gives the closed-form Dirichlet-kernel centroid;
                         transposing the phase swaps x and y; wfsCent.py's orientation asserts.
  add_power_spec_central (P) independent index-by-index loop (tiny sizes), clip and pad cases.
  prebin_image           (P) block sums computed with reshape (textbook).
  flip_array             (P) equals numpy.fft.fftshift for even sizes (mirror=1); involutions.
  prepare_spot_otf,      (P) direct circular convolution (brute force) on small arrays; a delta
  do_convolution         spot gives the identity times P^2.
  tot_sig                (D) with a spot: images of a clip = P = F run sum to sig * fracSubArea
                         (Parseval and the unnormalised r2c/c2r pair), 2-D and 4-D spots.
  bin_weights,           (P) integer ratio equals block sums; total flux preserved for every
  bin_image_literal,     ratio; equals bin_matrix (the separable operator) on random inputs.
  bin_matrix
  poisson_from_normal,   (P) exact inverse CDF below lambda=10 against pmf sums; mean/variance
  normal_cdf             and total-variation distance to the exact pmf above it.
  read_ccd               (D) floor ramps for types 0 and 1 (the pixel equal to the floor), read
                         noise sigma*z + readBg, unknown types untouched, the relative floor
                         option per image and its absolute and all-negative cases.
  select_brightest       (D) permutations (top N kept, minus the next brightest for -N), ties at
                         the threshold, N = 0 or >= npxl.
  apply_power            (D) powers 2, 1, 0 (no power), 0.5 and 3.
  optically_bin          (D) window column/row sums for ncen < nimg; for ncen = nimg the closed
                         form of the reference's in-place overwrite of row 1.
  ob_cents               (D) delta positions; a non-positive flux keeps the weighted sum.
  r2hc*/hc2r*,           (P) calc_correlation_hc equals correlate_fft (numpy ifft2 route,
  hc_multiply,           util/correlation.py:177-186) and brute-force circular correlation;
  transform_psf,         the half-complex unpack equals conj(fft2(fftshift(psf))).
  calc_correlation_hc,
  correlate_fft,
  hc_pattern_to_spectrum
  threshold_correlation  (D) thresholds 0, 0.3 (fraction of the maximum), 1 and 2 (absolute) on a
                         signed ramp; an all-negative image.
  diff_correlation       (P, D) modes 1 and 3: a pattern moved by a known shift gives its only
                         zero at that shift, constant images give the closed-form normalisation;
                         the fallback loop (correlationCentroiding 3): the same with its
                         opposite shift convention, and (n-|dy|)(n-|dx|) for constant images.
  brute_correlation      (P, D) zero shift is the sum of squares; two deltas correlate only at
                         the shift that brings them together.
  compute_cog            (P, D) delta images give the pixel offset exactly; symmetric images give
                         0; weighted: two deltas with weights, the ncen window, zero weighted flux.
  make_fit_mx,           (P) exact vertex recovery for sampled paraboloids; fit matrix equals the
  compute_fit            inverse of the normal matrix from an independent construction.
  gaussian_log           (P) exact centre of a sampled 2-D Gaussian through the Gaussian fit; the
                         returned image of a Gaussian-fit run equals the log of a plain run's.
  interp1                (D) knots, midpoints and clamping of a piecewise-linear table.
  apply_cal_poly         (D) polynomial identities ((1+c)^2, c^3, constants, empty).
  interpolate_phase      (P) reproduces linear and quadratic surfaces exactly away from the edges
                         (a property of the Keys cubic kernel the reference uses); edge samples.
  run_image              (D) subapLocation: the default tiling equals no locations; a smaller
                         window lands in the image's top-left corner, keeps the previous pixels
                         elsewhere and floors only the exposed ones.
  run_phase, _image_backend
                         compositions of the functions above; pinned through them and through
                         the whole-pipeline identities (flat phase, Parseval totals, ramps).
Functions without an independent check: none of the above; the exact FFTW/GSL arithmetic and
noise streams of the reference are outside the oracle (pre-drawn normals replace them).
"""
import functools
import math

import numpy as np

f64 = np.float64

# Update codes of the reference module (cmod/src/centmodule.c:109-130, util/centcmod.py:19-40).
CALSOURCE, SIG, ADDPOISSON, NCEN, READNOISE, READBG, NOISEFLOOR, SKYBRIGHTNESS, PXLPOWER = range(1, 10)
SPOTPSF, NTHREADS, OPTICALBINNING, CENTWEIGHT, CORRELATIONCENTROIDING, CORRTHRESH = range(10, 16)
CORRPATTERN, CALDATA, REFCENTS, CALCOEFF, USEBRIGHTEST, INTEGSTEPS, OUTPUTPHASEARR = range(16, 23)


# ----------------------------------------------------------------------------------------------
# image formation
# ----------------------------------------------------------------------------------------------
def tilt_fn(n, fftsize):
    """Half-pixel centring tilt, cmod/src/centmodule.c:3074-3081:
    t[i,j] = (pi/F)(F+1)(i+j+1-n)."""
    i = np.arange(n, dtype=f64)
    return (math.pi / fftsize * (fftsize + 1.0)) * (i[:, None] + i[None, :] + 1 - n)


def subap_windows(phs2d, nsubx, n):
    """(npup,npup) -> (nsubx*nsubx, n, n): sub-aperture s=sy*nsubx+sx reads
    phs[(sy*n+i)*npup + sx*n+j] (cmod/src/centmodule.c:479-486)."""
    a = np.asarray(phs2d)
    return a.reshape(nsubx, n, nsubx, n).transpose(0, 2, 1, 3).reshape(nsubx * nsubx, n, n)


def add_power_spec_central(power, fftsize, psfsize):
    """|X|^2 of the central psfsize window, rotated 180 degrees, zero-padded when psfsize>F.
    cmod/src/centmodule.c:430-471 (additive==0 branch). power: (S,F,F)."""
    S = power.shape[0]
    hll = np.zeros((S, psfsize, psfsize), f64)
    if psfsize > fftsize:
        rowstart, rowoff, c = (psfsize - fftsize) // 2, 0, fftsize
    else:
        rowstart, rowoff, c = 0, (fftsize - psfsize) // 2, psfsize
    sub = power[:, rowoff:rowoff + c, rowoff:rowoff + c]
    hll[:, rowstart:rowstart + c, rowstart:rowstart + c] = sub[:, ::-1, ::-1]
    return hll


def compute_hll(phs_windows, pupfn, fftsize, paddedsize, amp_windows=None):
    """computeHll, cmod/src/centmodule.c:473-509: E = pup*exp(i(phi-t)) zero-padded to F x F,
    unnormalised forward FFT (FFTW_FORWARD, :3105), then add_power_spec_central."""
    S, n, _ = phs_windows.shape
    t = tilt_fn(n, fftsize)
    arg = phs_windows.astype(f64) - t[None]
    E = pupfn.astype(f64) * np.exp(1j * arg)
    if amp_windows is not None:
        E = E * amp_windows.astype(f64)
    X = np.fft.fft2(E, s=(fftsize, fftsize))
    power = X.real ** 2 + X.imag ** 2
    return add_power_spec_central(power, fftsize, paddedsize)


def prebin_image(hll, paddedsize, pbf, psfsize):
    """prebinImage, cmod/src/centmodule.c:1583-1626: bin by pbf from paddedsize, clip or
    zero-pad (centred) to psfsize."""
    S = hll.shape[0]
    scaled = paddedsize // pbf
    binned = scaled
    offset = 0
    if scaled > psfsize:
        offset = (scaled - psfsize) // 2
        binned = psfsize
    out = np.zeros((S, binned, binned), f64)
    for i in range(binned):
        for j in range(binned):
            r0 = i * pbf + offset
            c0 = j * pbf + offset
            out[:, i, j] = hll[:, r0:r0 + pbf, c0:c0 + pbf].sum(axis=(1, 2))
    if scaled < psfsize:
        off = (psfsize - scaled) // 2
        res = np.zeros((S, psfsize, psfsize), f64)
        res[:, off:off + scaled, off:off + scaled] = out
        return res
    return out


def _flip_quad(a, r0, c0, h, w):
    """flipQuad, cmod/src/centmodule.c:253-274 (rotate one h x w block by 180 degrees)."""
    blk = a[r0:r0 + h, c0:c0 + w].copy()
    a[r0:r0 + h, c0:c0 + w] = blk[::-1, ::-1]


def flip_array(arr, mirror):
    """flipArray, cmod/src/centmodule.c:275-349 on a 2-D square array (returns a copy).
    mirror==1 swaps the quadrants (fftshift for even sizes); mirror==0 flips each quadrant."""
    a = np.array(arr, dtype=f64, copy=True)
    n = a.shape[0]
    if mirror == 0:
        w1 = n // 2
        w3 = (n + 1) // 2
        h1 = w1
        _flip_quad(a, 0, 0, h1, w1)
        _flip_quad(a, h1, 0, w3, w1)
        _flip_quad(a, 0, w1, h1, w3)
        _flip_quad(a, h1, w1, w3, w3)
        return a
    if n % 2 == 0:
        f2 = n // 2
        out = np.empty_like(a)
        out[:f2, :f2] = a[f2:, f2:]
        out[f2:, f2:] = a[:f2, :f2]
        out[:f2, f2:] = a[f2:, :f2]
        out[f2:, :f2] = a[:f2, f2:]
        return out
    b = flip_array(a, 0)
    return b[::-1, ::-1].copy()


def prepare_spot_otf(spotpsf):
    """prepareSpotPsf, cmod/src/centmodule.c:2303-2375: OTF = r2c(flipArray(spot, mirror=1)).
    spotpsf: (P,P) or (S,P,P). Returns complex (.., P, P//2+1)."""
    sp = np.asarray(spotpsf, f64)
    if sp.ndim == 2:
        return np.fft.rfft2(flip_array(sp, 1))
    sp = sp.reshape(-1, sp.shape[-2], sp.shape[-1])
    return np.stack([np.fft.rfft2(flip_array(s, 1)) for s in sp])


def do_convolution(hll, otf, psfsize, clipsize):
    """doConvolution, cmod/src/centmodule.c:524-564: hll <- c2r(r2c(hll)*OTF) (FFTW c2r is
    unnormalised, so x P^2) then central clip to clipsize."""
    F = np.fft.rfft2(hll, s=(psfsize, psfsize))
    conv = np.fft.irfft2(F * otf, s=(psfsize, psfsize)) * (psfsize * psfsize)
    if clipsize < psfsize:
        off = (psfsize - clipsize) // 2
        conv = conv[:, off:off + clipsize, off:off + clipsize]
    return conv


def bin_weights(clipsize, nimg):
    """prepareBinImage, cmod/src/centmodule.c:565-608 (thispxl, nextpxl, indx).
    Weights are float in the reference; here float64 of the same ratios."""
    thispxl = np.zeros(clipsize, f64)
    nextpxl = np.zeros(clipsize, f64)
    indx = np.zeros(clipsize, np.int64)
    cnt = 0
    for i in range(clipsize):
        cnt += nimg
        if cnt >= clipsize:
            excess = cnt - clipsize
            thispxl[i] = (nimg - excess) / float(nimg)
            nextpxl[i] = excess / float(nimg)
            indx[i] = 1
            cnt = excess
        else:
            thispxl[i] = 1.0
            nextpxl[i] = 0.0
            indx[i] = 0
    return thispxl, nextpxl, indx


def bin_image_literal(img, clipsize, nimg):
    """binImage, cmod/src/centmodule.c:609-651, loop for loop (vectorised over sub-apertures).
    Note the reference applies nextpxl[j] to input pixel j+1 (and nextpxl[i] to row i+1)."""
    img = np.asarray(img, f64)
    if clipsize == nimg:
        return img.copy()
    S = img.shape[0]
    out = np.zeros((S, nimg, nimg), f64)
    t, nx, idx = bin_weights(clipsize, nimg)
    ii = 0
    for i in range(clipsize - 1):
        jj = 0
        for j in range(clipsize - 1):
            out[:, ii, jj] += t[i] * t[j] * img[:, i, j]
            if jj + 1 < nimg:
                out[:, ii, jj + 1] += t[i] * nx[j] * img[:, i, j + 1]
            if ii + 1 < nimg:
                out[:, ii + 1, jj] += nx[i] * t[j] * img[:, i + 1, j]
                if jj + 1 < nimg:
                    out[:, ii + 1, jj + 1] += nx[i] * nx[j] * img[:, i + 1, j + 1]
            jj += idx[j]
        j = clipsize - 1
        out[:, ii, jj] += t[i] * t[j] * img[:, i, j]
        if ii + 1 < nimg:
            out[:, ii + 1, jj] += nx[i] * t[j] * img[:, i + 1, j]
        ii += idx[i]
    i = clipsize - 1
    jj = 0
    for j in range(clipsize - 1):
        out[:, ii, jj] += t[i] * t[j] * img[:, i, j]
        if jj + 1 < nimg:
            out[:, ii, jj + 1] += t[i] * nx[j] * img[:, i, j + 1]
        jj += idx[j]
    j = clipsize - 1
    out[:, ii, jj] += t[i] * t[j] * img[:, i, j]
    return out


def bin_matrix(clipsize, nimg):
    """Separable form A (nimg x clipsize) of binImage: img_out = A @ img @ A.T.
    Derived from cmod/src/centmodule.c:609-651: input row r reaches output row ii(r) with weight
    thispxl[r] and output row ii(r-1)+1 with weight nextpxl[r-1] (if that row exists); the last
    input row/column has no 'next' term of its own."""
    if clipsize == nimg:
        return np.eye(nimg)
    t, nx, idx = bin_weights(clipsize, nimg)
    A = np.zeros((nimg, clipsize), f64)
    pos = np.zeros(clipsize, np.int64)
    ii = 0
    for r in range(clipsize):
        pos[r] = ii
        ii += idx[r]
    for r in range(clipsize):
        A[pos[r], r] += t[r]
        if r >= 1 and pos[r - 1] + 1 < nimg:
            A[pos[r - 1] + 1, r] += nx[r - 1]
    return A


def tot_sig(pupfn, fftsize, spotpsf=None):
    """updateTotSigArr, cmod/src/centmodule.c:2403-2427: sum(pup)*F^2 [* sum(spot)*P^2]."""
    S = pupfn.shape[0]
    tot = pupfn.reshape(S, -1).astype(f64).sum(axis=1) * fftsize * fftsize
    if spotpsf is not None:
        sp = np.asarray(spotpsf, f64)
        P = sp.shape[-1]
        if sp.ndim == 2:
            tot = tot * sp.sum() * P * P
        else:
            tot = tot * sp.reshape(S, -1).sum(axis=1) * P * P
    return tot


# ----------------------------------------------------------------------------------------------
# noise (pre-drawn arrays replace the in-module RNGs; see poisson_from_normal)
# ----------------------------------------------------------------------------------------------
POISSON_SMALL = 10.0      # same split as rk_poisson (cmod/src/centmodule.c:760-782)
POISSON_SMALL_KMAX = 40   # cap of the exact summation branch (P(X>=40 | lam<10) < 1e-11)


def normal_cdf(z):
    return 0.5 * np.vectorize(math.erfc)(-np.asarray(z, f64) / math.sqrt(2.0))


def poisson_from_normal(lam, z):
    """Deterministic Poisson(lam) variate from a pre-drawn standard normal z.

    Replaces rk_poisson (cmod/src/centmodule.c:760-782), which draws a variable number of PCG32
    uniforms per pixel. lam<=0 gives 0 (as rk_poisson). 0<lam<10: exact inverse CDF of
    u=Phi(z) by summing the pmf from k=0 (the search stops when the sum no longer grows or at
    k=40). lam>=10: third-order Cornish-Fisher quantile y of the smoothed lattice variable
    Y = X + U(-1/2,1/2), whose cumulants are k1=lam, k2=lam+1/12, k3=lam, k4=lam-1/120:
        s = sqrt(lam+1/12), g1 = lam/s^3, g2 = (lam-1/120)/s^4,
        w = z + g1 (z^2-1)/6 + g2 (z^3-3z)/24 - g1^2 (2z^3-5z)/36,
        k = max(ceil(lam + s w - 1/2), 0).
    Its total-variation distance to the exact Poisson pmf is about 0.041/lam (4.1e-3 at lam=10,
    1.7e-4 at lam=300).
    """
    lam = np.asarray(lam, f64)
    z = np.asarray(z, f64)
    out = np.zeros(np.broadcast(lam, z).shape, f64)
    lam_b = np.broadcast_to(lam, out.shape)
    z_b = np.broadcast_to(z, out.shape)
    small = (lam_b > 0) & (lam_b < POISSON_SMALL)
    if small.any():
        lm = lam_b[small]
        u = normal_cdf(z_b[small])
        k = np.zeros_like(lm)
        p = np.exp(-lm)
        F = p.copy()
        active = u > F
        for kk in range(1, POISSON_SMALL_KMAX + 1):
            if not active.any():
                break
            p = np.where(active, p * lm / kk, p)
            Fn = np.where(active, F + p, F)
            grew = Fn > F
            k = np.where(active & grew, kk, k)
            active = active & grew & (u > Fn)
            F = Fn
        out[small] = k
    big = lam_b >= POISSON_SMALL
    if big.any():
        lm = lam_b[big]
        zz = z_b[big]
        s = np.sqrt(lm + 1.0 / 12.0)
        g1 = lm / (s * s * s)
        g2 = (lm - 1.0 / 120.0) / (s * s * s * s)
        z2 = zz * zz
        w = zz + g1 * (z2 - 1.0) / 6.0 + g2 * (z2 * zz - 3.0 * zz) / 24.0 - g1 * g1 * (2.0 * z2 * zz - 5.0 * zz) / 36.0
        out[big] = np.maximum(np.ceil(lm + s * w - 0.5), 0.0)
    return out


def read_ccd(img, zread, readbg, readnoise, floor, thresh_type, relative_floor=False):
    """readCCD, cmod/src/centmodule.c:785-817 with the Gaussian draw replaced by readnoise*zread
    (zread pre-drawn N(0,1)); readnoise==0 adds readbg only. thresh_type 0: x<f?0:x-f;
    1: x<f?0:x; any other type: the data are left untouched (no noise, no readbg, no floor).
    relative_floor (engine option, not in the reference; thresh_type 0 only): the floor of each
    image (first axis) is f*max(x) when 0<f<=1 (thresholdCorrelation's rule, :1488-1505)."""
    x = np.asarray(img, f64)
    if thresh_type not in (0, 1):
        return x   # the reference prints "Unknown threshold type" and leaves the data (:797, :813)
    if readnoise == 0 or zread is None:
        x = x + readbg
    else:
        x = x + (readnoise * np.asarray(zread, f64) + readbg)
    if thresh_type == 1:
        return np.where(x < floor, 0.0, x)
    if not relative_floor:
        return np.where(x < floor, 0.0, x - floor)
    S = x.shape[0]
    if 0 < floor <= 1:
        thr = floor * np.maximum(x.reshape(S, -1).max(axis=1), 0.0)
    else:
        thr = np.full(S, float(floor))
    thr = thr.reshape((S,) + (1,) * (x.ndim - 1))
    return np.where(x < thr, 0.0, x - thr)


def select_brightest(img, use_brightest):
    """selectBrightestPixels, cmod/src/centmodule.c:1544-1581 on one flat image (copy)."""
    x = np.array(img, f64).ravel()
    npxl = x.size
    ub = int(use_brightest)
    subtract = False
    if ub < 0:
        ub = -ub
        subtract = True
    if 0 < ub < npxl:
        srt = np.sort(x)
        thr = srt[npxl - ub]
        if subtract:
            k = npxl - ub - 1
            while k >= 0 and srt[k] == thr:
                k -= 1
            sub = srt[k] if k >= 0 else 0.0
            x = np.where(x < thr, 0.0, x - sub)
        else:
            x = np.where(x < thr, 0.0, x)
    return x.reshape(np.shape(img))


def apply_power(img, p):
    """cmod/src/centmodule.c:1893-1897: p==2 squares, p not in {0,1} uses pow."""
    if p == 2:
        return img * img
    if p != 1 and p != 0:
        return np.power(img, p)
    return img


def optically_bin(img, nimg, ncen):
    """opticallyBin, cmod/src/centmodule.c:819-850 on one (nimg,nimg) image (copy, flat)."""
    a = np.array(img, f64).ravel()
    s = (nimg - ncen) // 2
    e = nimg - s
    v0 = sum(a[nimg * s + i] for i in range(s, e))
    if s != 0:
        a[:ncen] = 0
    for i in range(s, e):
        for j in range(1 if s == 0 else s, e):
            a[i - s] += a[j * nimg + i]
    for i in range(s + 1, e):
        if s != 0:
            a[i - s + ncen] = 0
        for j in range(1 if s == 0 else s, e):
            a[i - s + ncen] += a[i * nimg + j]
    a[ncen] = v0
    a[2 * ncen:] = 0
    return a.reshape(nimg, nimg)


def ob_cents(img, ncen):
    """computeOBCents, cmod/src/centmodule.c:852-872. Without positive flux in a direction the
    reference leaves that direction's weighted sum in place (it is not reset to 0)."""
    a = np.asarray(img, f64).ravel()
    i = np.arange(ncen)
    sx = a[:ncen].sum()
    sy = a[ncen:2 * ncen].sum()
    cx = (a[:ncen] * i).sum()
    cy = (a[ncen:2 * ncen] * i).sum()
    if sx > 0:
        cx = cx / sx - (ncen - 1) / 2.0
    if sy > 0:
        cy = cy / sy - (ncen - 1) / 2.0
    return cx, cy


# ----------------------------------------------------------------------------------------------
# correlation
# ----------------------------------------------------------------------------------------------
def r2hc1d(a):
    """FFTW R2HC of the last axis: r0..r_{n/2}, i_{(n+1)/2-1}..i_1 (util/correlation.py:63-74)."""
    a = np.asarray(a, f64)
    n = a.shape[-1]
    res = np.fft.fft(a, axis=-1)
    b = np.empty_like(a)
    b[..., :n // 2 + 1] = res.real[..., :n // 2 + 1]
    if n // 2 + 1 < n:
        b[..., n // 2 + 1:] = res.imag[..., (n + 1) // 2 - 1:0:-1]
    return b


def hc2r1d(a):
    """FFTW HC2R (unnormalised inverse) of the last axis."""
    a = np.asarray(a, f64)
    n = a.shape[-1]
    spec = np.zeros(a.shape, np.complex128)
    spec.real[..., :n // 2 + 1] = a[..., :n // 2 + 1]
    for k in range(1, (n + 1) // 2):
        spec.imag[..., k] = a[..., n - k]
    for k in range(n // 2 + 1, n):
        spec[..., k] = np.conj(spec[..., n - k])
    return np.fft.ifft(spec, axis=-1).real * n


def r2hc2d(a):
    """fftwf_plan_r2r_2d(R2HC,R2HC): separable R2HC over both axes."""
    return np.swapaxes(r2hc1d(np.swapaxes(r2hc1d(a), -1, -2)), -1, -2)


def hc2r2d(a):
    return np.swapaxes(hc2r1d(np.swapaxes(hc2r1d(a), -1, -2)), -1, -2)


def transform_psf(psf):
    """util/correlation.py:20-51: r2hc(fftshift(psf)) with the conjugating sign flips, float32."""
    p = np.asarray(psf, f64)

    def one(x):
        ff = r2hc2d(np.fft.fftshift(x))
        mm = ff.shape[0] // 2 + 1
        nn = ff.shape[1] // 2 + 1
        ff[:mm, nn:] *= -1
        ff[mm:, :nn] *= -1
        return ff
    if p.ndim == 2:
        return one(p).astype(np.float32)
    out = np.empty_like(p)
    for idx in np.ndindex(p.shape[:-2]):
        out[idx] = one(p[idx])
    return out.astype(np.float32)


def hc_multiply(a, B):
    """The half-complex product loop of calcCorrelation, cmod/src/centmodule.c:1436-1482, on
    one (m,n) array pair (returns a new array)."""
    a = np.array(a, f64, copy=True)
    B = np.asarray(B, f64)
    m, n = a.shape
    a[0, 0] *= B[0, 0]
    neven = n % 2 == 0
    meven = m % 2 == 0
    if neven:
        a[0, n // 2] *= B[0, n // 2]
    if meven:
        a[m // 2, 0] *= B[m // 2, 0]
        if neven:
            a[m // 2, n // 2] *= B[m // 2, n // 2]
    for i in range(1, (n + 1) // 2):
        r1 = a[0, i] * B[0, i] - a[0, n - i] * B[0, n - i]
        r2 = a[0, i] * B[0, n - i] + a[0, n - i] * B[0, i]
        a[0, i] = r1
        a[0, n - i] = r2
        if meven:
            r3 = a[m // 2, i] * B[m // 2, i] - a[m // 2, n - i] * B[m // 2, n - i]
            r4 = a[m // 2, i] * B[m // 2, n - i] + a[m // 2, n - i] * B[m // 2, i]
            a[m // 2, i] = r3
            a[m // 2, n - i] = r4
    for i in range(1, (m + 1) // 2):
        r5 = a[i, 0] * B[i, 0] - a[m - i, 0] * B[m - i, 0]
        r6 = a[i, 0] * B[m - i, 0] + a[m - i, 0] * B[i, 0]
        a[i, 0] = r5
        a[m - i, 0] = r6
        if neven:
            r7 = a[i, n // 2] * B[i, n // 2] - a[m - i, n // 2] * B[m - i, n // 2]
            r8 = a[i, n // 2] * B[m - i, n // 2] + a[m - i, n // 2] * B[i, n // 2]
            a[i, n // 2] = r7
            a[m - i, n // 2] = r8
        for j in range(1, (n + 1) // 2):
            aij, amn, ain, amj = a[i, j], a[m - i, n - j], a[i, n - j], a[m - i, j]
            r1 = aij * B[i, j] + amn * B[m - i, n - j] - ain * B[i, n - j] - amj * B[m - i, j]
            r2 = aij * B[m - i, n - j] + amn * B[i, j] + amj * B[i, n - j] + ain * B[m - i, j]
            r3 = aij * B[i, n - j] - amn * B[m - i, j] + ain * B[i, j] - amj * B[m - i, n - j]
            r4 = aij * B[m - i, j] - amn * B[i, n - j] + amj * B[i, j] - ain * B[m - i, n - j]
            a[i, j] = r1
            a[m - i, n - j] = r2
            a[i, n - j] = r3
            a[m - i, j] = r4
    return a


def calc_correlation_hc(img, pattern_hc, nimg, corrsize):
    """calcCorrelation, cmod/src/centmodule.c:1403-1486: centre-pad img to corrsize, R2HC 2-D,
    half-complex multiply by the pre-transformed pattern, HC2R (unnormalised). One subap."""
    tmp = np.zeros((corrsize, corrsize), f64)
    o = (corrsize - nimg) // 2
    tmp[o:o + nimg, o:o + nimg] = img
    a = r2hc2d(tmp)
    a = hc_multiply(a, pattern_hc)
    return hc2r2d(a)


def correlate_fft(img_padded, ref):
    """util/correlation.py:181: ifft2(fft2(a)*conj(fft2(fftshift(b)))).real, scaled by m*n to
    match FFTW's unnormalised inverse."""
    m, n = img_padded.shape[-2:]
    return np.fft.ifft2(np.fft.fft2(img_padded) *
                        np.conj(np.fft.fft2(np.fft.fftshift(ref, axes=(-2, -1))))).real * m * n


def hc_pattern_to_spectrum(pattern_hc):
    """Full complex spectrum K with calc_correlation_hc(img) == m*n*ifft2(fft2(pad(img))*K).real.
    Obtained by applying the reference's HC product to unit impulses in the HC domain."""
    p = np.asarray(pattern_hc, f64)
    m, n = p.shape
    # the HC product is linear in a; for a real input x, R2HC(x) -> hc_multiply -> HC2R equals a
    # circular correlation whose kernel is hc2r2d(hc_multiply(r2hc2d(delta_00), p)).
    delta = np.zeros((m, n))
    delta[0, 0] = 1.0
    kern = hc2r2d(hc_multiply(r2hc2d(delta), p))   # response to a unit impulse at (0,0)
    return np.fft.fft2(kern) / (m * n)


def threshold_correlation(img, thresh):
    """thresholdCorrelation, cmod/src/centmodule.c:1488-1505 on one image (copy)."""
    x = np.asarray(img, f64)
    t = thresh
    if 0 < thresh <= 1.0:
        mx = max(0.0, x.max())
        t = mx * thresh
    return np.where(x < t, 0.0, x - t)


def _corr_windows(nimg, corrsize, ncen, i, j):
    """Shared index arithmetic of diffCorrelation/bruteCorrelation (:1226-1260, :1361-1386)."""
    miny = minx = corrsize if corrsize < nimg else nimg
    half = _cdiv(nimg - corrsize, 2)
    offy = i - ncen // 2
    my = miny
    srefy = max(-offy - half, 0)
    siny = max(half + offy, 0)
    if siny + my > nimg:
        my = nimg - siny
    if srefy + my > corrsize:
        my = corrsize - srefy
    offx = j - ncen // 2
    mx = minx
    srefx = max(-offx - half, 0)
    sinx = max(half + offx, 0)
    if sinx + mx > nimg:
        mx = nimg - sinx
    if srefx + mx > corrsize:
        mx = corrsize - srefx
    return srefy, siny, my, srefx, sinx, mx


def _cdiv(a, b):
    """C integer division (truncates toward zero)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def diff_correlation(bimg, pattern, nimg, corrsize, ncen, mode):
    """diffCorrelation, cmod/src/centmodule.c:1203-1338 for mode 1 (sum d^2 / npxl) and mode 3
    (sum d^2 / (flux*flux)); other modes follow the reference's fallback loop (:1309-1336)."""
    out = np.zeros((corrsize, corrsize), f64)
    b = np.asarray(bimg, f64)
    c = np.asarray(pattern, f64)
    if mode in (1, 3):
        for i in range(ncen):
            for j in range(ncen):
                srefy, siny, my, srefx, sinx, mx = _corr_windows(nimg, corrsize, ncen, i, j)
                if my <= 0 or mx <= 0:
                    s = 0.0
                    cc = c[0:0, 0:0]
                    bb = b[0:0, 0:0]
                else:
                    cc = c[srefy:srefy + my, srefx:srefx + mx]
                    bb = b[siny:siny + my, sinx:sinx + mx]
                d = cc - bb
                s = float((d * d).sum())
                oi = i + (corrsize - ncen) // 2
                oj = j + (corrsize - ncen) // 2
                if mode == 1:
                    if mx != 0 and my != 0:
                        out[oi, oj] = s / (mx * my)
                else:
                    tot = float(bb.sum())
                    tot2 = float(cc.sum())
                    if tot != 0 and tot2 != 0:
                        out[oi, oj] = s / (tot * tot2)
        return out
    for i in range(ncen):
        my = corrsize - abs(ncen // 2 - i)
        for j in range(ncen):
            mx = corrsize - abs(ncen // 2 - j)
            s = 0.0
            for y in range(my):
                for x in range(mx):
                    if i < ncen // 2:
                        if j < ncen // 2:
                            d = c[y, x] - b[ncen // 2 - i + y, ncen // 2 - j + x]
                        else:
                            d = c[y, j - ncen // 2 + x] - b[ncen // 2 - i + y, x]
                    else:
                        if j < ncen // 2:
                            d = c[i - ncen // 2 + y, x] - b[y, ncen // 2 - j + x]
                        else:
                            d = c[i - ncen // 2 + y, j - ncen // 2 + x] - b[y, x]
                    s += abs(d)
            out[i + (nimg - ncen) // 2, j + (nimg - ncen) // 2] = s * s / (my * mx)
    return out


def brute_correlation(bimg, pattern, nimg, corrsize, ncen):
    """bruteCorrelation, cmod/src/centmodule.c:1339-1398."""
    out = np.zeros((corrsize, corrsize), f64)
    b = np.asarray(bimg, f64)
    c = np.asarray(pattern, f64)
    for i in range(ncen):
        for j in range(ncen):
            srefy, siny, my, srefx, sinx, mx = _corr_windows(nimg, corrsize, ncen, i, j)
            s = 0.0
            if my > 0 and mx > 0:
                s = float((c[srefy:srefy + my, srefx:srefx + mx] *
                           b[siny:siny + my, sinx:sinx + mx]).sum())
            if mx != 0 and my != 0:
                out[i + (corrsize - ncen) // 2, j + (corrsize - ncen) // 2] = s
    return out


# ----------------------------------------------------------------------------------------------
# centroiding
# ----------------------------------------------------------------------------------------------
def compute_cog(img, nimg, ncen, weighting=None, centre_on_brightest=0):
    """computeCoG, cmod/src/centmodule.c:873-971 on one (nimg,nimg) image.
    Returns (cx, cy): x = column direction, y = row direction."""
    a = np.asarray(img, f64).reshape(nimg, nimg)
    if centre_on_brightest == 0:
        sx = (nimg - ncen) // 2
        ex = nimg - sx
        sy, ey = sx, ex
    else:
        flat = a.ravel()
        if centre_on_brightest == 1:
            mx = 0.0
            pos = 0
            for i in range(nimg * nimg):
                if flat[i] > mx:
                    mx = flat[i]
                    pos = i
        else:
            mx = 0.0
            pos = 0
            for i in range(nimg * nimg):
                if flat[i] < mx or mx == 0:
                    mx = flat[i]
                    pos = i
        sx = int(math.trunc(pos % nimg - ncen / 2.0 + 0.5))
        sy = int(math.trunc(pos // nimg - ncen / 2.0 + 0.5))
        sx = min(max(sx, 0), nimg - ncen)
        sy = min(max(sy, 0), nimg - ncen)
        ex = sx + ncen
        ey = sy + ncen
    win = a[sy:ey, sx:ex]
    if weighting is not None:
        win = win * np.asarray(weighting, f64).reshape(nimg, nimg)[sy:ey, sx:ex]
    jj = np.arange(sx, ex, dtype=f64)[None, :]
    ii = np.arange(sy, ey, dtype=f64)[:, None]
    s = win.sum()
    if weighting is None and centre_on_brightest not in (0, 1):
        cx = (-win * jj).sum()
        cy = (-win * ii).sum()
    else:
        cx = (win * jj).sum()
        cy = (win * ii).sum()
    if s > 0:
        cx /= s
        cy /= s
        if centre_on_brightest != 0:
            cx += sx
            cy += sy
        return cx - (nimg - 1) / 2.0, cy - (nimg - 1) / 2.0
    return 0.0, 0.0


@functools.lru_cache(maxsize=None)
def _fit_mx_cached(fitsize):
    r = make_fit_mx_uncached(fitsize)
    r.setflags(write=False)
    return r


def make_fit_mx(fitsize):
    """makeFitMx (see make_fit_mx_uncached), computed once per fit size."""
    return _fit_mx_cached(int(fitsize))


def make_fit_mx_uncached(fitsize):
    """makeFitMx, cmod/src/centmodule.c:997-1057: first 5 rows of the inverse of the 6x6 normal
    matrix of the basis [x^2, xy, y^2, x, y, 1] over a fitsize x fitsize grid."""
    y, x = np.mgrid[0:fitsize, 0:fitsize].astype(f64)
    basis = np.stack([x * x, x * y, y * y, x, y, np.ones_like(x)]).reshape(6, -1)
    mx = basis @ basis.T
    return np.linalg.inv(mx)[:5]


def gaussian_log(img, gmin, grep):
    """The in-place transform of computeFit's Gaussian branch, cmod/src/centmodule.c:1087-1098
    (logf in FP32 there)."""
    x = np.asarray(img, f64)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(x < gmin, grep, np.log(np.where(x < gmin, 1.0, x)))


def compute_fit(img, nimg, ncen, fitsize, do_gaussian=0, gmin=0.0, grep=0.0, corr_mode=0):
    """computeFit, cmod/src/centmodule.c:1069-1188 on one (nimg,nimg) image (nimg is the
    reference's corrsize). Returns (cx, cy)."""
    a = np.array(img, f64).ravel()
    nc = (nimg - ncen) // 2
    quad = False
    if fitsize < 0:
        quad = True
        fitsize = 3
    if fitsize < 3:
        fitsize = 3
    ssum = 0.0
    mx = 0.0
    mxpos = 0
    if do_gaussian:
        for i in range(nimg * nimg):
            ssum += a[i]
            if a[i] > mx:
                mxpos = i
                mx = a[i]
            a[i] = grep if a[i] < gmin else math.log(a[i])
    elif corr_mode in (0, 1, 5):
        for i in range(nimg * nimg):
            ssum += a[i]
            if a[i] > mx:
                mxpos = i
                mx = a[i]
    else:
        mx = a[nc * nimg + nc]
        mxpos = nc * nimg + nc
        for i in range(nc, nc + ncen):
            for j in range(nc, nc + ncen):
                ssum += a[i * nimg + j]
                if a[i * nimg + j] < mx:
                    mxpos = i * nimg + j
                    mx = a[i * nimg + j]
    if not ssum >= 0:
        return 0.0, 0.0
    startx = int(mxpos % nimg - fitsize // 2)
    starty = int(mxpos // nimg - fitsize // 2)
    startx = max(startx, nc)
    starty = max(starty, nc)
    startx = min(startx, nc + ncen - fitsize)
    starty = min(starty, nc + ncen - fitsize)
    img2 = a.reshape(nimg, nimg)
    if not quad:
        win = img2[starty:starty + fitsize, startx:startx + fitsize]
        yy, xx = np.mgrid[0:fitsize, 0:fitsize].astype(f64)
        vec = np.array([(win * xx * xx).sum(), (win * xx * yy).sum(), (win * yy * yy).sum(),
                        (win * xx).sum(), (win * yy).sum(), win.sum()])
        res = make_fit_mx(fitsize) @ vec
        if do_gaussian:
            cx = -res[3] / (2 * res[0])
            cy = -res[4] / (2 * res[2])
        else:
            cx = (res[1] * res[4] / (2 * res[2]) - res[3]) / (2.0 * res[0] - res[1] * res[1] / (2.0 * res[2]))
            cy = -(res[4] + res[1] * cx) / (2.0 * res[2])
    else:
        w = img2[starty:starty + 3, startx:startx + 3].ravel()
        n3 = 3
        a2 = (w[2 * n3 + 1] - w[1]) / 2.0
        a3 = (w[2 * n3 + 1] - 2 * w[n3 + 1] + w[1]) / 2.0
        a4 = (w[n3 + 2] - w[n3]) / 2.0
        a5 = (w[n3 + 2] - 2 * w[n3 + 1] + w[n3]) / 2.0
        a6 = (w[2 * n3 + 2] - w[2] - w[2 * n3] + w[0]) / 4.0
        cy = (2 * a2 * a5 - a4 * a6) / (a6 * a6 - 4 * a3 * a5) + 1
        cx = (2 * a3 * a4 - a2 * a6) / (a6 * a6 - 4 * a3 * a5) + 1
    if not (cx >= 0 and cy >= 0 and cx < fitsize and cy < fitsize):
        cx = mxpos % nimg - startx
        cy = mxpos // nimg - starty
    cy -= ncen / 2.0 - 0.5
    cx -= ncen / 2.0 - 0.5
    cx += startx + ncen / 2.0 - nimg / 2.0
    cy += starty + ncen / 2.0 - nimg / 2.0
    return cx, cy


def interp1(x, xp, fp):
    """interp, cmod/src/centmodule.c:1506-1519 (n = len(xp)-1)."""
    n = len(xp) - 1
    if x <= xp[0]:
        return fp[0]
    if x >= xp[n]:
        return fp[n]
    i = 1
    while x > xp[i]:
        i += 1
    return (x - xp[i - 1]) * (fp[i] - fp[i - 1]) / (xp[i] - xp[i - 1]) + fp[i - 1]


def apply_cal_poly(c, coeff):
    """applyCentroidCalibrationInterpolation, cmod/src/centmodule.c:1520-1530."""
    res = 0.0
    xc = 1.0
    for k in range(len(coeff)):
        res += xc * coeff[k]
        xc *= c
    return res


# ----------------------------------------------------------------------------------------------
# bicubic phase interpolation (npup != nsubx*n)
# ----------------------------------------------------------------------------------------------
_KEYS = np.array([0, 2, 0, 0, -1, 0, 1, 0, 2, -5, 4, -1, -1, 3, -3, 1], f64).reshape(4, 4)


def interpolate_phase(inp, nsubx, n):
    """interpolatePhase, cmod/src/centmodule.c:1628-1758 for all sub-apertures: bicubic (Keys)
    resampling of the (m,m) input onto the (nsubx*n)^2 grid spanning the same extent."""
    inp = np.asarray(inp, f64)
    m = inp.shape[0]
    tot = nsubx * n
    scale = (m - 1.0) / (tot - 1)
    out = np.zeros((tot, tot), f64)

    def coords(o):
        t = scale * o
        i0 = int(math.floor(t))
        t -= i0
        i0 -= 1
        if o == tot - 1:
            t = 1.0
            i0 = m - 3
        return i0, t

    def fetch(r, c):
        r = min(max(r, 0), m - 1)
        c = min(max(c, 0), m - 1)
        return inp[r, c]
    for y in range(tot):
        iny, ty = coords(y)
        tv = np.array([0.5, 0.5 * ty, 0.5 * ty * ty, 0.5 * ty ** 3])
        for x in range(tot):
            inx, tx = coords(x)
            tt = np.array([0.5, 0.5 * tx, 0.5 * tx * tx, 0.5 * tx ** 3])
            p = np.zeros(4)
            for i in range(4):
                if iny < 0:
                    f1 = [fetch(r, inx + i) for r in (0, 1, 2)]
                    f = [f1[0], f1[0], f1[1], f1[2]]
                elif iny + 4 > m:
                    f1 = [fetch(r, inx + i) for r in (m - 3, m - 2, m - 1)]
                    f = [f1[0], f1[1], f1[2], f1[2]]
                else:
                    f = [fetch(iny + r, inx + i) for r in range(4)]
                p[i] = tv @ (_KEYS @ np.array(f))
            out[y, x] = tt @ (_KEYS @ p)
    return out


# ----------------------------------------------------------------------------------------------
# whole pipeline
# ----------------------------------------------------------------------------------------------
class Params:
    """The scalar and array arguments of cent.initialise (cmod/src/centmodule.c:2876-2883,
    util/centcmod.py:115), plus the state later set through update (refCents, calibration).
    Arrays are in the reference layout; pupfn is (nsubaps, n, n)."""

    def __init__(self, nsubx, n, fftsize, clipsize, nimg, ncen, pupfn, subflag, fracSubArea,
                 **kw):
        self.nsubx = nsubx
        self.nsubaps = nsubx * nsubx
        self.phasesize = n
        self.fftsize = fftsize
        self.clipsize = clipsize
        self.nimg = nimg
        self.ncen = ncen
        self.pupfn = np.asarray(pupfn, f64).reshape(self.nsubaps, n, n)
        self.subflag = np.asarray(subflag).reshape(-1).astype(np.int64)
        self.fracSubArea = np.asarray(fracSubArea, f64).reshape(-1)
        d = dict(readnoise=0.0, readbg=0.0, addPoisson=0, noiseFloor=0.0, sig=1.0, sigArr=None,
                 skybrightness=0.0, skybrightnessArr=None, calsource=0, pxlPower=1.0,
                 spotpsf=None, opticalBinning=0, centWeight=None, correlationCentroiding=0,
                 corrThresh=0.0, corrPattern=None, threshType=0, imageOnly=0, useBrightest=0,
                 useBrightestArr=None, preBinningFactor=1, parabolicFit=0, gaussianFit=0,
                 gaussianMinVal=0.0, gaussianReplaceVal=0.0, refCents=None, calData=None,
                 calBounds=None, calSteps=None, calCoeff=None, relativeFloor=False)
        for k, v in kw.items():
            if k not in d:
                raise TypeError("unknown parameter %s" % k)
            d[k] = v
        self.__dict__.update(d)

    # derived sizes: setSpotPsf (cmod/src/centmodule.c:2255-2301)
    def sizes(self):
        if self.spotpsf is None:
            return self.clipsize, self.clipsize, 1          # psfsize, paddedsize, prebin
        P = np.asarray(self.spotpsf).shape[-1]
        pbf = self.preBinningFactor
        padded = P if pbf <= 1 else self.fftsize
        return P, padded, max(pbf, 1)

    def corrsize(self):
        if self.correlationCentroiding and self.corrPattern is not None:
            return np.asarray(self.corrPattern).shape[-1]
        return self.nimg


def _as_f64(p):
    """A copy of the parameters with the per-subaperture arrays the backend reads converted to
    float64 once (the backend indexes them per subaperture)."""
    q = Params.__new__(Params)
    q.__dict__.update(p.__dict__)
    for name in ("corrPattern", "centWeight", "calData", "calSteps", "calCoeff", "refCents"):
        v = getattr(p, name)
        if v is not None:
            setattr(q, name, np.asarray(v, f64))
    for name in ("useBrightestArr", "calBounds"):
        v = getattr(p, name)
        if v is not None:
            setattr(q, name, np.asarray(v))
    return q


def _image_backend(p, img, s, nexposed_full):
    """Steps after the noise (brightest, power, centroid, calibration, refCents) for one subap;
    shared by the phase and the image paths (:1883-1961, :2032-2110)."""
    nimg = p.nimg
    flat = np.array(img, f64).ravel()
    if p.calsource == 0:
        ub = p.useBrightest if p.useBrightestArr is None else int(np.asarray(p.useBrightestArr).ravel()[s])
        flat[:nexposed_full] = select_brightest(flat[:nexposed_full], ub)
    flat[:nexposed_full] = apply_power(flat[:nexposed_full], p.pxlPower)
    x = flat.reshape(nimg, nimg)
    out_img = x
    corr_out = None
    if p.imageOnly:
        return out_img, None, corr_out
    cs = p.corrsize()
    cc = p.correlationCentroiding
    pat = None if p.corrPattern is None else np.asarray(p.corrPattern, f64).reshape(-1, cs, cs)[s]
    if cc == 1 and pat is not None:
        corr = calc_correlation_hc(x, pat, nimg, cs)
        corr = threshold_correlation(corr, p.corrThresh)
        corr_out = corr
        src, size = corr, cs
    elif 1 < cc < 5 and pat is not None:
        corr = diff_correlation(x, pat, nimg, cs, p.ncen, cc - 1)
        corr = threshold_correlation(corr, p.corrThresh)
        corr_out = corr
        src, size = corr, cs
    elif cc == 5 and pat is not None:
        corr = brute_correlation(x, pat, nimg, cs, p.ncen)
        corr = threshold_correlation(corr, p.corrThresh)
        corr_out = corr
        src, size = corr, cs
    else:
        src, size = x, cs
    if p.opticalBinning == 1:
        cx, cy = ob_cents(x, p.ncen)
    elif p.parabolicFit != 0 or p.gaussianFit > 0:
        cx, cy = compute_fit(src, size, p.ncen, p.parabolicFit + p.gaussianFit, p.gaussianFit,
                             p.gaussianMinVal, p.gaussianReplaceVal, cc)
        if p.gaussianFit:
            # computeFit takes the logarithm in place (:1087-1098) on the buffer it is given: the
            # correlation image when correlating (:1899-1906), otherwise the output image (:1919)
            g = gaussian_log(src, p.gaussianMinVal, p.gaussianReplaceVal)
            if corr_out is not None:
                corr_out = g
            else:
                out_img = g
    else:
        w = None
        if p.centWeight is not None:
            cw = np.asarray(p.centWeight, f64)
            w = cw if cw.ndim == 2 else cw.reshape(-1, size, size)[s]
        cob = 1 if cc == 5 else cc
        cx, cy = compute_cog(src, size, p.ncen, w, cob)
    if p.calData is not None:
        cd = np.asarray(p.calData, f64).reshape(2, p.nsubaps, -1)
        cb = np.asarray(p.calBounds).reshape(2, p.nsubaps, 2)
        st = np.asarray(p.calSteps, f64)
        res = []
        for d, c in ((0, cx), (1, cy)):
            lo, hi = int(cb[d, s, 0]), int(cb[d, s, 1])
            res.append(interp1(c, cd[d, s, lo:hi + 1], st[lo:hi + 1]))
        cx, cy = res
    elif p.calCoeff is not None:
        co = np.asarray(p.calCoeff, f64).reshape(p.nsubaps, 2, -1)
        cx = apply_cal_poly(cx, co[s, 0])
        cy = apply_cal_poly(cy, co[s, 1])
    if p.refCents is not None:
        rc = np.asarray(p.refCents, f64).reshape(-1)
        cx -= rc[2 * s]
        cy -= rc[2 * s + 1]
    return out_img, (cx, cy), corr_out


def run_phase(p, phs, zpois=None, zread=None, bimg_in=None):
    """centroidsFromPhase, cmod/src/centmodule.c:1760-1984, one readout (curInteg cycling is
    handled by the caller). phs: (npup,npup) phase, or (2,npup,npup) phase+amplitude, or a
    grid of another size (bicubic interpolation, :1791-1805).
    zpois, zread: pre-drawn N(0,1) arrays of shape (nsubaps, nimg, nimg) replacing the module's
    Poisson and Gaussian generators. Returns (cents (nsubaps,2), bimg (nsubaps,nimg,nimg),
    corrimg (nsubaps,corrsize,corrsize) or None)."""
    nsubx, n, F = p.nsubx, p.phasesize, p.fftsize
    S = p.nsubaps
    nimg = p.nimg
    phs = np.asarray(phs, f64)
    amp = None
    if phs.ndim == 3:
        phs, amp = phs[0], phs[1]
    if phs.shape[0] != nsubx * n:
        phs = interpolate_phase(phs, nsubx, n)
        if amp is not None:
            amp = interpolate_phase(amp, nsubx, n)
    win = subap_windows(phs, nsubx, n)
    awin = None if amp is None else subap_windows(amp, nsubx, n)
    P, padded, pbf = p.sizes()
    hll = compute_hll(win, p.pupfn, F, padded, awin)
    if pbf > 1:
        hll = prebin_image(hll, padded, pbf, P)
    if p.spotpsf is not None:
        otf = prepare_spot_otf(p.spotpsf)
        if otf.ndim == 2:
            otf = otf[None]
        hll = do_convolution(hll, otf, P, p.clipsize)
    img = bin_image_literal(hll, p.clipsize, nimg)
    totsig = tot_sig(p.pupfn, F, p.spotpsf)
    if awin is not None:
        totpup = p.pupfn.reshape(S, -1).sum(axis=1)
        tsa = (p.pupfn * awin).reshape(S, -1).sum(axis=1)
        totsig = np.where(totpup != 0, tsa * totsig / np.where(totpup != 0, totpup, 1), 0.0)
    sigv = (np.asarray(p.sigArr, f64).ravel() if p.sigArr is not None else p.sig * p.fracSubArea)
    scale = np.where(totsig > 0, sigv / np.where(totsig > 0, totsig, 1.0), 1.0)
    img = img * scale[:, None, None]
    sky = p.skybrightness * p.fracSubArea
    img = img + np.where(sky > 0, sky, 0.0)[:, None, None]
    if p.skybrightnessArr is not None:
        img = img + np.asarray(p.skybrightnessArr, f64).reshape(S, nimg, nimg)
    cents = np.zeros((S, 2), f64)
    p = _as_f64(p)
    bimg = np.zeros((S, nimg, nimg), f64) if bimg_in is None else np.array(bimg_in, f64).reshape(S, nimg, nimg)
    cs = p.corrsize()
    corrimg = None
    if p.correlationCentroiding and p.corrPattern is not None:
        corrimg = np.zeros((S, cs, cs), f64)
    for s in range(S):
        if p.subflag[s] != 1:
            if p.imageOnly:
                rn = p.readnoise * (p.calsource == 0)
                z = None if zread is None else zread[s]
                bimg[s] = read_ccd(np.zeros((1, nimg, nimg)), None if z is None else z[None],
                                   p.readbg, rn, p.noiseFloor, p.threshType, p.relativeFloor)[0]
            continue
        x = img[s]
        if p.opticalBinning == 1:
            x = optically_bin(x, nimg, p.ncen)
            nexp = 2 * p.ncen
        else:
            nexp = nimg * nimg
        if (p.addPoisson == 1 and p.calsource == 0) and (
                totsig[s] > 0 or sky[s] > 0 or p.skybrightnessArr is not None):
            flat = x.ravel().copy()
            flat[:nexp] = poisson_from_normal(flat[:nexp], np.asarray(zpois[s], f64).ravel()[:nexp])
            x = flat.reshape(nimg, nimg)
        rn = p.readnoise * (p.calsource == 0)
        flat = x.ravel().copy()
        zr = None if zread is None else np.asarray(zread[s], f64).ravel()[:nexp][None]
        flat[:nexp] = read_ccd(flat[:nexp][None], zr, p.readbg, rn, p.noiseFloor, p.threshType,
                               p.relativeFloor)[0]
        x = flat.reshape(nimg, nimg)
        out_img, c, corr = _image_backend(p, x, s, nexp)
        bimg[s] = out_img
        if c is not None:
            cents[s] = c
        if corr is not None:
            corrimg[s] = corr
    return cents, bimg, corrimg


def run_image(p, image, bimg_in=None, subap_location=None, zread=None):
    """centroidsFromImage, cmod/src/centmodule.c:1987-2127: image (nsubx*nimg)^2 2-D or
    (nsubx,nsubx,nimg,nimg) 4-D."""
    S = p.nsubaps
    nimg = p.nimg
    image = np.asarray(image, f64)
    p = _as_f64(p)
    cents = np.zeros((S, 2), f64)
    bimg = np.zeros((S, nimg, nimg), f64) if bimg_in is None else np.array(bimg_in, f64).reshape(S, nimg, nimg)
    cs = p.corrsize()
    corrimg = None
    if p.correlationCentroiding and p.corrPattern is not None:
        corrimg = np.zeros((S, cs, cs), f64)
    for s in range(S):
        if p.subflag[s] != 1:
            if p.imageOnly:
                rn = p.readnoise * (p.calsource == 0)
                z = None if zread is None else np.asarray(zread[s], f64)[None]
                bimg[s] = read_ccd(np.zeros((1, nimg, nimg)), z, p.readbg, rn,
                                   p.noiseFloor, p.threshType, p.relativeFloor)[0]
            continue
        x = bimg[s].copy()
        if image.ndim == 2:
            if subap_location is None:
                r0 = (s // p.nsubx) * nimg
                c0 = (s % p.nsubx) * nimg
                r1, c1 = r0 + nimg, c0 + nimg
            else:
                r0, r1, c0, c1 = [int(v) for v in np.asarray(subap_location).reshape(-1, 4)[s]]
            h, w = r1 - r0, c1 - c0
            flat = x.ravel()
            for j in range(h):
                for k in range(w):
                    flat[j * nimg + k] = image[r0 + j, c0 + k]
            x = flat.reshape(nimg, nimg)
            nexp = h * w
        else:
            x = image.reshape(S, nimg, nimg)[s].copy()
            nexp = nimg * nimg
        if p.calsource == 0:
            flat = x.ravel().copy()
            flat[:nexp] = read_ccd(flat[:nexp][None], None, 0.0, 0.0, p.noiseFloor, p.threshType,
                                   p.relativeFloor)[0]
            x = flat.reshape(nimg, nimg)
        out_img, c, corr = _image_backend(p, x, s, nexp)
        bimg[s] = out_img
        if c is not None:
            cents[s] = c
        if corr is not None:
            corrimg[s] = corr
    return cents, bimg, corrimg
