"""Pins of the oracle's detector, correlation-threshold, calibration and image-input functions.

Each ``pin_*`` checks one oracle function against closed forms that follow from the reference's
C text (cmod/src/centmodule.c) or from mathematics, not against a retyped copy of the oracle's own
formula. ``test_pin_*`` runs it on the oracle; ``test_mutant_fails_its_pin`` plants one plausible
mistake at a time in a copy of the oracle function (a comparison turned from < to <=, a dropped
term, a window off by one, a wrong shift direction) and requires the pin to reject it.
"""
import inspect
import textwrap

import numpy as np
import pytest

from oracle import centoracle as O

aeq = np.testing.assert_array_equal
aclose = np.testing.assert_allclose


# ---- readCCD (:785-817) ----------------------------------------------------------------------
def pin_read_ccd(read_ccd):
    x = np.linspace(-1.0, 3.0, 17)[None]          # steps of 0.25: x + bg hits the floor exactly
    bg, f = 0.25, 1.0
    v = x + bg
    # type 0: x<f ? 0 : x-f  ==  max(x - f, 0)
    aeq(read_ccd(x, None, bg, 0.0, f, 0), np.maximum(v - f, 0.0))
    # type 1: x<f ? 0 : x  -- the pixel equal to the floor is kept
    assert (v == f).any()
    aeq(read_ccd(x, None, bg, 0.0, f, 1), np.where(v >= f, v, 0.0))
    # read noise: sigma * z + readBg, floor far below the data
    z = np.linspace(-2.0, 2.0, 17)[None]
    aclose(read_ccd(x, z, bg, 0.5, -10.0, 1), x + 0.5 * z + bg, rtol=0, atol=1e-15)
    aclose(read_ccd(x, z, bg, 0.5, -10.0, 0), x + 0.5 * z + bg + 10.0, rtol=0, atol=1e-14)
    aeq(read_ccd(x, z, bg, 0.0, -10.0, 1), v)      # readNoise == 0: no draw at all
    # unknown threshold types: "not adding read noise", data untouched (:797, :813)
    for t in (2, 3, -1):
        aeq(read_ccd(x, z, bg, 0.5, f, t), x)
    # relative floor (engine option): per-image threshold f * max(image)
    imgs = np.stack([np.linspace(0.0, 4.0, 9), np.linspace(-1.0, 2.0, 9)])
    aclose(read_ccd(imgs, None, 0.0, 0.0, 0.25, 0, True),
           np.stack([np.maximum(imgs[0] - 1.0, 0), np.maximum(imgs[1] - 0.5, 0)]), rtol=0, atol=1e-15)
    aeq(read_ccd(imgs, None, 0.0, 0.0, 1.5, 0, True), np.maximum(imgs - 1.5, 0))   # > 1: absolute
    aeq(read_ccd(-np.abs(imgs) - 1, None, 0.0, 0.0, 0.5, 0, True), np.zeros_like(imgs))   # max <= 0


# ---- thresholdCorrelation (:1488-1505) -------------------------------------------------------
def pin_threshold_correlation(tc):
    x = np.linspace(-2.0, 6.0, 33).reshape(3, 11)   # max 6, includes 0 and 1.8 exactly
    aeq(tc(x, 0.0), np.where(x < 0, 0.0, x))
    aclose(tc(x, 0.3), np.where(x < 1.8, 0.0, x - 1.8), rtol=0, atol=1e-15)
    aeq(tc(x, 2.0), np.where(x < 2.0, 0.0, x - 2.0))   # above 1: an absolute threshold
    aeq(tc(x, 1.0), np.zeros_like(x))                 # 1: the maximum itself
    aeq(tc(-np.abs(x) - 1.0, 0.5), np.zeros_like(x))  # no positive value: threshold 0


# ---- selectBrightestPixels (:1544-1581) ------------------------------------------------------
def pin_select_brightest(sb):
    rng = np.random.default_rng(3)
    perm = rng.permutation(16).astype(float)          # distinct values 0..15
    aeq(sb(perm, 4), np.where(perm >= 12, perm, 0.0))
    aeq(sb(perm, -4), np.where(perm >= 12, perm - 11, 0.0))   # minus the next brightest (11)
    for ub in (0, 16, 20, -16):
        aeq(sb(perm, ub), perm)                        # 0 or >= npxl: untouched
    ties = np.array([5.0, 1.0, 5.0, 3.0, 5.0, 0.0, 2.0, 3.0])
    aeq(sb(ties, 2), np.where(ties >= 5, ties, 0.0))             # every pixel equal to thr is kept
    aeq(sb(ties, -2), np.where(ties >= 5, ties - 3.0, 0.0))      # next value below the ties: 3
    aeq(sb(np.full(6, 2.0), -3), np.full(6, 2.0))                # nothing below: subtract 0
    aeq(sb(perm.reshape(4, 4), 3), np.where(perm >= 13, perm, 0.0).reshape(4, 4))


# ---- opticallyBin / computeOBCents (:819-872) ------------------------------------------------
def pin_optically_bin(ob):
    rng = np.random.default_rng(5)
    nimg = 8
    for ncen in (4, 6):                               # s = (nimg - ncen)/2 > 0: plain window sums
        img = rng.random((nimg, nimg))
        s, e = (nimg - ncen) // 2, nimg - (nimg - ncen) // 2
        out = ob(img, nimg, ncen).ravel()
        win = img[s:e, s:e]
        aclose(out[:ncen], win.sum(axis=0), rtol=1e-14)          # x: column sums of the window
        aclose(out[ncen:2 * ncen], win.sum(axis=1), rtol=1e-14)  # y: row sums (row s first)
        aeq(out[2 * ncen:], 0.0)
    # s == 0 (ncen == nimg): the y sums are written into row 1 while it is still being read, so
    # row i's sum takes img[1, i] in place of its column-0 term (row 0's sum is exact)
    img = rng.random((nimg, nimg))
    out = ob(img, nimg, nimg).ravel()
    aclose(out[:nimg], img.sum(axis=0), rtol=1e-14)
    want_y = np.array([img[0].sum()] + [img[1, i] + img[i, 1:].sum() for i in range(1, nimg)])
    aclose(out[nimg:2 * nimg], want_y, rtol=1e-14)
    aeq(out[2 * nimg:], 0.0)


def pin_ob_cents(oc):
    ncen = 6
    for k in range(ncen):
        a = np.zeros(4 * ncen)
        a[k] = 2.0
        a[ncen + (ncen - 1 - k)] = 3.0
        cx, cy = oc(a, ncen)
        assert cx == k - (ncen - 1) / 2.0 and cy == (ncen - 1 - k) - (ncen - 1) / 2.0
    a = np.zeros(2 * ncen)
    a[[1, 4]] = [-3.0, 1.0]                            # no positive flux in x: weighted sum stays
    a[ncen + 2] = 5.0
    cx, cy = oc(a, ncen)
    assert cx == -3.0 * 1 + 1.0 * 4 and cy == 2 - (ncen - 1) / 2.0
    assert oc(np.zeros(2 * ncen), ncen) == (0.0, 0.0)


# ---- calibration (:1506-1530) -----------------------------------------------------------------
def pin_interp1(f):
    xp, fp = np.array([0.0, 1.0, 3.0, 6.0]), np.array([10.0, 12.0, 8.0, 8.0])
    for x, want in [(-5, 10), (0, 10), (1, 12), (3, 8), (6, 8), (9, 8), (0.5, 11), (2, 10), (2.5, 9), (4.5, 8), (6.5, 8)]:
        assert abs(f(x, xp, fp) - want) < 1e-12, (x, f(x, xp, fp), want)


def pin_apply_cal_poly(f):
    for c in (-1.5, 0.0, 0.25, 2.0):
        assert abs(f(c, np.array([1.0, 2.0, 1.0])) - (1 + c) ** 2) < 1e-12
        assert abs(f(c, np.array([0.0, 0.0, 0.0, 1.0])) - c ** 3) < 1e-12
        assert f(c, np.array([7.0])) == 7.0
    assert f(3.0, np.array([])) == 0.0


# ---- pixel power (:1893-1897, applyPower :517-522) -------------------------------------------
def pin_apply_power(ap):
    x = np.array([[0.0, 0.5, 1.0, 2.0, 9.0]])
    aeq(ap(x, 2.0), x * x)
    aeq(ap(x, 1.0), x)
    aeq(ap(x, 0.0), x)                                 # 0 means "no power", not x**0
    aclose(ap(x, 0.5), np.sqrt(x), rtol=1e-15)
    aclose(ap(x, 3.0), x * x * x, rtol=1e-15)


# ---- weighted computeCoG (:873-971) -----------------------------------------------------------
def pin_weighted_cog(cog):
    n = 8
    img = np.zeros((n, n))
    img[2, 5], img[6, 1] = 2.0, 1.0
    w = np.ones((n, n))
    w[2, 5], w[6, 1] = 0.5, 3.0                        # weighted fluxes 1 and 3
    cx, cy = cog(img, n, n, w)
    aclose([cx, cy], [(1 * 5 + 3 * 1) / 4 - 3.5, (1 * 2 + 3 * 6) / 4 - 3.5], rtol=0, atol=1e-14)
    cx, cy = cog(img, n, 4, w)                         # ncen 4: window [2, 6)^2 keeps (2, 5) only
    aclose([cx, cy], [5 - 3.5, 2 - 3.5], rtol=0, atol=1e-14)
    w4 = np.zeros((n, n))
    assert cog(img, n, n, w4) == (0.0, 0.0)            # zero weighted flux


# ---- diffCorrelation, fallback loop (correlationCentroiding 3, :1309-1336) ---------------------
def pin_diff_correlation_loop(dc):
    rng = np.random.default_rng(11)
    for nimg, ncen in ((8, 8), (8, 6), (10, 7)):
        h = ncen // 2
        off = (nimg - ncen) // 2
        c = rng.random((nimg, nimg))
        for ry, rx in ((0, 0), (1, -2), (-1, 1), (2, 2)):
            if abs(ry) > h or abs(rx) > h or h - ry >= ncen or h - rx >= ncen:
                continue
            b = np.roll(c, (ry, rx), axis=(0, 1))      # b[Z] = c[Z - r]
            out = dc(b, c, nimg, nimg, ncen, 2)
            win = out[off:off + ncen, off:off + ncen]
            i, j = np.unravel_index(np.argmin(win), win.shape)
            assert (i, j) == (h - ry, h - rx) and win[i, j] == 0.0, (nimg, ncen, ry, rx, i, j)
            assert (np.delete(win.ravel(), i * ncen + j) > 0).all()
        # pattern 1, image 0: sum |d| = overlap area, so each entry is (n - |dy|)(n - |dx|)
        out = dc(np.zeros((nimg, nimg)), np.ones((nimg, nimg)), nimg, nimg, ncen, 2)
        k = np.arange(ncen)
        want = np.zeros((nimg, nimg))
        want[off:off + ncen, off:off + ncen] = np.outer(nimg - abs(h - k), nimg - abs(h - k))
        aeq(out, want)


# ---- diffCorrelation modes 1 and 3 (:1226-1306) and bruteCorrelation (:1339-1398) -------------
def pin_diff_correlation_modes(dc):
    rng = np.random.default_rng(12)
    for nimg, ncen in ((8, 8), (8, 6), (10, 7)):
        h, off = ncen // 2, (nimg - ncen) // 2
        c = rng.random((nimg, nimg)) + 0.5
        for mode in (1, 3):
            for ry, rx in ((0, 0), (1, -2), (-2, 1)):
                if abs(ry) > h or abs(rx) > h or h + ry >= ncen or h + rx >= ncen:
                    continue
                b = np.roll(c, (ry, rx), axis=(0, 1))  # the image is the pattern moved by r
                win = dc(b, c, nimg, nimg, ncen, mode)[off:off + ncen, off:off + ncen]
                i, j = np.unravel_index(np.argmin(win), win.shape)
                assert (i, j) == (h + ry, h + rx) and win[i, j] == 0.0, (mode, nimg, ncen, ry, rx, i, j)
        # pattern 2, image 1: d = 1 on the whole overlap, so mode 1 gives mean d^2 = 1 and mode 3
        # area / (sum(image) * sum(pattern)) = area / (area * 2 area)
        k = np.arange(ncen)
        area = np.outer(nimg - abs(k - h), nimg - abs(k - h))
        for mode, want in ((1, np.ones((ncen, ncen))), (3, 1.0 / (2.0 * area))):
            out = dc(np.ones((nimg, nimg)), np.full((nimg, nimg), 2.0), nimg, nimg, ncen, mode)
            aclose(out[off:off + ncen, off:off + ncen], want, rtol=1e-14)
            out[off:off + ncen, off:off + ncen] = 0
            aeq(out, 0.0)


def pin_brute_correlation(bc):
    nimg = 8
    for ncen in (8, 5):
        h, off = ncen // 2, (nimg - ncen) // 2
        for (py, px), (qy, qx) in (((3, 4), (3, 4)), ((5, 2), (4, 3)), ((1, 6), (3, 5))):
            b = np.zeros((nimg, nimg))
            c = np.zeros((nimg, nimg))
            b[py, px], c[qy, qx] = 2.0, 3.0
            out = bc(b, c, nimg, nimg, ncen)
            want = np.zeros((nimg, nimg))
            i, j = h + py - qy, h + px - qx               # the shift that brings the deltas together
            if 0 <= i < ncen and 0 <= j < ncen:
                want[off + i, off + j] = 6.0
            aeq(out, want)


# ---- centroidsFromImage with subapLocation (:2009-2030) ----------------------------------------
def pin_run_image_subap_location(run_image):
    nsubx, nimg = 2, 4
    rng = np.random.default_rng(2)
    image = rng.random((nsubx * nimg, nsubx * nimg)) + 0.1
    p = O.Params(nsubx, 4, 8, 8, nimg, nimg, np.ones((4, 4, 4)), np.ones(4, np.int64), np.ones(4))
    tiles = np.array([[(s // 2) * 4, (s // 2) * 4 + 4, (s % 2) * 4, (s % 2) * 4 + 4] for s in range(4)], np.int32)
    c0, b0, _ = run_image(p, image)
    c1, b1, _ = run_image(p, image, subap_location=tiles)
    aeq(c0, c1)
    aeq(b0, b1)
    for s in range(4):                                # the default cut-out is the s-th tile
        aeq(b0[s], image[tiles[s, 0]:tiles[s, 1], tiles[s, 2]:tiles[s, 3]])
    loc = tiles.copy()
    loc[3] = [1, 4, 5, 7]                             # a 3 x 2 window of subaperture 3
    prev = np.full((4, nimg, nimg), -7.0)
    c2, b2, _ = run_image(p, image, bimg_in=prev, subap_location=loc)
    want = prev[3].copy()
    want[:3, :2] = image[1:4, 5:7]                    # top-left corner; the rest keeps bimg_in
    want.ravel()[:6] = np.maximum(want.ravel()[:6], 0.0)   # floor 0 on the exposed pixels only
    aeq(b2[3], want)
    aclose(c2[3], O.compute_cog(want, nimg, nimg), rtol=0, atol=1e-14)
    aeq(c2[:3], c0[:3])


# ---- totsig with a spot kernel (:2403-2427) ----------------------------------------------------
def pin_tot_sig_spot(tot_sig):
    """With a binary pupil, clip = psfsize = fftsize and no binning, nothing is clipped: the images
    sum to sig * fracSubArea exactly when totsig = sum(pup) F^2 sum(spot) P^2 (Parseval and the
    unnormalised r2c/c2r pair of doConvolution)."""
    nsubx, n, F = 2, 4, 8
    rng = np.random.default_rng(9)
    pup = np.ones((4, n, n))
    pup[0, 0, :2] = 0.0
    pup[3, 1:3, 1:3] = 0.0
    for spot in (rng.random((F, F)), rng.random((nsubx, nsubx, F, F))):
        sp = np.asarray(spot)
        orig = O.tot_sig
        O.tot_sig = tot_sig
        try:
            p = O.Params(nsubx, n, F, F, F, F, pup, np.ones(4, np.int64), pup.reshape(4, -1).mean(axis=1),
                         sig=7.0, spotpsf=sp)
            _, b, _ = O.run_phase(p, rng.standard_normal((nsubx * n, nsubx * n)))
        finally:
            O.tot_sig = orig
        aclose(b.reshape(4, -1).sum(axis=1), 7.0 * pup.reshape(4, -1).mean(axis=1), rtol=1e-12)


PINS = {
    "read_ccd": (pin_read_ccd, lambda: O.read_ccd),
    "threshold_correlation": (pin_threshold_correlation, lambda: O.threshold_correlation),
    "select_brightest": (pin_select_brightest, lambda: O.select_brightest),
    "optically_bin": (pin_optically_bin, lambda: O.optically_bin),
    "ob_cents": (pin_ob_cents, lambda: O.ob_cents),
    "interp1": (pin_interp1, lambda: O.interp1),
    "apply_cal_poly": (pin_apply_cal_poly, lambda: O.apply_cal_poly),
    "compute_cog": (pin_weighted_cog, lambda: O.compute_cog),
    "apply_power": (pin_apply_power, lambda: O.apply_power),
    "diff_correlation": (pin_diff_correlation_loop, lambda: O.diff_correlation),
    "run_image": (pin_run_image_subap_location, lambda: O.run_image),
    "diff_correlation_modes": (pin_diff_correlation_modes, lambda: O.diff_correlation),
    "brute_correlation": (pin_brute_correlation, lambda: O.brute_correlation),
    "tot_sig": (pin_tot_sig_spot, lambda: O.tot_sig),
}


@pytest.mark.parametrize("name", sorted(PINS))
def test_pin(name):
    pin, fn = PINS[name]
    pin(fn())


def _mutant(fn, old, new, helper=None):
    """A copy of oracle function fn with the first occurrence of old replaced by new, in fn itself
    or, when helper is given, in that oracle helper (fn then calls the mutated helper)."""
    target = helper or fn
    src = textwrap.dedent(inspect.getsource(target))
    assert old in src, (target.__name__, old)
    ns = dict(vars(O))
    exec(compile(src.replace(old, new, 1), "<mutant>", "exec"), ns)
    if helper is not None:
        exec(compile(textwrap.dedent(inspect.getsource(fn)), "<caller>", "exec"), ns)
    return ns[fn.__name__]


MUTANTS = [
    ("read_ccd", "np.where(x < floor, 0.0, x)", "np.where(x <= floor, 0.0, x)"),
    ("read_ccd", "np.where(x < floor, 0.0, x - floor)", "np.where(x < floor, 0.0, x)"),
    ("read_ccd", "x = x + readbg", "x = x + 0.0"),
    ("read_ccd", "if thresh_type not in (0, 1):", "if thresh_type not in (0, 1, 2):"),
    ("read_ccd", "x.reshape(S, -1).max(axis=1)", "x.max() + 0 * x.reshape(S, -1).max(axis=1)"),
    ("threshold_correlation", "if 0 < thresh <= 1.0:", "if 0 < thresh < 1.0:"),
    ("threshold_correlation", "t = mx * thresh", "t = thresh"),
    ("select_brightest", "thr = srt[npxl - ub]", "thr = srt[npxl - ub - 1]"),
    ("select_brightest", "while k >= 0 and srt[k] == thr:", "while False:"),
    ("select_brightest", "subtract = True", "subtract = False"),
    ("optically_bin", "for j in range(1 if s == 0 else s, e):", "for j in range(s, e):"),
    ("optically_bin", "a[ncen] = v0", "pass"),
    ("optically_bin", "v0 = sum(a[nimg * s + i] for i in range(s, e))", "v0 = sum(a[nimg * s + i] for i in range(s, e - 1))"),
    ("ob_cents", "cx = cx / sx - (ncen - 1) / 2.0", "cx = cx / sx - ncen / 2.0"),
    ("ob_cents", "if sx > 0:", "if sx >= 0:"),
    ("interp1", "return (x - xp[i - 1]) * (fp[i] - fp[i - 1]) / (xp[i] - xp[i - 1]) + fp[i - 1]",
     "return (x - xp[i - 1]) * (fp[i] - fp[i - 1]) / (xp[i] - xp[i - 1]) + fp[i]"),
    ("interp1", "if x >= xp[n]:", "if x > xp[n] + 1:"),
    ("apply_cal_poly", "xc *= c", "xc *= c + 0.0 * xc + (k == 0)"),
    ("apply_power", "if p != 1 and p != 0:", "if p != 1:"),
    ("apply_power", "return img * img", "return np.power(img, 2.0) + 0 * img[..., :1] + (img > 8)"),
    ("compute_cog", "win = win * np.asarray(weighting, f64).reshape(nimg, nimg)[sy:ey, sx:ex]", "win = win"),
    ("compute_cog", "win = a[sy:ey, sx:ex]", "win = a[sy:ey + 1, sx:ex + 1]"),
    ("diff_correlation", "d = c[y, x] - b[ncen // 2 - i + y, ncen // 2 - j + x]",
     "d = c[y, x] - b[ncen // 2 - i + y, x]"),
    ("diff_correlation", "s * s / (my * mx)", "s / (my * mx)"),
    ("diff_correlation", "d = c[i - ncen // 2 + y, j - ncen // 2 + x] - b[y, x]",
     "d = c[i - ncen // 2 + y, j - ncen // 2 + x] - b[y + (y > 0), x]"),
    ("diff_correlation_modes", "srefy = max(-offy - half, 0)", "srefy = max(offy - half, 0)", "_corr_windows"),
    ("diff_correlation_modes", "if srefx + mx > corrsize:", "if srefx + mx > corrsize + 1:", "_corr_windows"),
    ("diff_correlation_modes", "out[oi, oj] = s / (tot * tot2)", "out[oi, oj] = s / (tot * tot)"),
    ("diff_correlation_modes", "out[oi, oj] = s / (mx * my)", "out[oi, oj] = s / (mx * mx)"),
    ("brute_correlation", "b[siny:siny + my, sinx:sinx + mx]", "b[srefy:srefy + my, sinx:sinx + mx]"),
    ("brute_correlation", "out[i + (corrsize - ncen) // 2", "out[i + (corrsize - ncen + 1) // 2"),
    ("run_image", "flat[j * nimg + k] = image[r0 + j, c0 + k]", "flat[j * w + k] = image[r0 + j, c0 + k]"),
    ("run_image", "nexp = h * w", "nexp = nimg * nimg"),
    ("tot_sig", "tot = tot * sp.sum() * P * P", "tot = tot * sp.sum() * P"),
    ("tot_sig", "tot = tot * sp.reshape(S, -1).sum(axis=1) * P * P", "tot = tot * sp.sum() / S * P * P"),
]


@pytest.mark.filterwarnings("ignore::RuntimeWarning")
@pytest.mark.parametrize("mut", MUTANTS, ids=["%s:%d" % (m[0], i) for i, m in enumerate(MUTANTS)])
def test_mutant_fails_its_pin(mut):
    name, old, new = mut[:3]
    pin, fn = PINS[name]
    bad = _mutant(fn(), old, new, getattr(O, mut[3]) if len(mut) > 3 else None)
    with pytest.raises((AssertionError, IndexError, ValueError)):
        pin(bad)
