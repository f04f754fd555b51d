"""CPU: pins of the atmosphere / DM oracle (oracle/atmosoracle.py) against closed forms and an
independent spline implementation, each with a planted mistake that the pin must reject."""
import numpy as np
import pytest

from oracle import atmosoracle as A


def test_linearshift_reproduces_bilinear_surfaces():
    """Bilinear interpolation is exact on a + b x + c y + d x y (x: column, y: row)."""
    ii, jj = np.mgrid[0:9, 0:11].astype(np.float64)
    a, b, c, d = 0.3, -1.25, 2.5, 0.125
    f = a + b * jj + c * ii + d * ii * jj
    for xs, ys in ((0.25, 0.5), (0.0, 0.0), (0.75, 0.125), (1.0, 1.0)):
        out = A.linearshift(f, xs, ys, 8, 10)
        x = jj[:8, :10] + np.float64(np.float32(xs))
        y = ii[:8, :10] + np.float64(np.float32(ys))
        # the weights are floats: exact up to their rounding
        want = a + b * x + c * y + d * x * y
        assert np.abs(out - want).max() < 1e-6 * np.abs(want).max()
    # zero shift is the window itself, a whole shift the next window
    assert np.array_equal(A.linearshift(f, 0, 0, 8, 10), f[:8, :10])
    assert np.array_equal(A.linearshift(f, 1, 1, 8, 10), f[1:9, 1:11])
    # planted mistake: the shifts swapped (x weighting the next row) fails on a column ramp
    ramp = jj.copy()
    s1, s2, s3, s4 = (np.float64(s) for s in A.linearshift_weights(0.25, 0.0))
    wrong = ramp[:8, :10] * s1 + ramp[:8, 1:11] * s2 + ramp[1:9, :10] * s3 + ramp[1:9, 1:11] * s4
    assert not np.allclose(wrong, A.linearshift(ramp, 0.25, 0.0, 8, 10))


def test_layer_window_shift_cases():
    """util/atmos.py:697-720 by hand: add-at-start layers move by 1 - interp, add-at-end by interp
    (or a whole pixel when interp is 0)."""
    assert A.layer_window_shift(3.0, 1, 0.25) == (3, 0.75)
    assert A.layer_window_shift(3.0, 1, 0.0) == (3, 0.0)
    assert A.layer_window_shift(3.0, -1, 0.25) == (3, 0.25)
    assert A.layer_window_shift(3.0, -1, 0.0) == (4, 0.0)
    w, f = A.layer_window_shift(2.6, -1, 0.5)
    assert w == 3 and abs(f - 0.1) < 1e-12


def test_make_layer_equals_a_longer_screen():
    """Moving a screen by (nr, nc) and inserting the next columns then rows equals cutting the next
    window out of one longer screen (every add direction)."""
    rng = np.random.default_rng(3)
    big = rng.standard_normal((30, 34))
    H, W, nr, nc = 20, 24, 2, 3
    for col_add in (-1, 1):
        for row_add in (-1, 1):
            r0 = 5 if row_add < 0 else 7
            c0 = 4 if col_add < 0 else 6
            scr = big[r0:r0 + H, c0:c0 + W]
            r1 = r0 + nr if row_add < 0 else r0 - nr
            c1 = c0 + nc if col_add < 0 else c0 - nc
            want = big[r1:r1 + H, c1:c1 + W]
            cols = (want[:, W - nc:] if col_add < 0 else want[:, :nc]).T
            rows = want[H - nr:] if row_add < 0 else want[:nr]
            got = A.make_layer(scr, cols, rows, col_add, row_add)
            assert np.array_equal(got, want), (col_add, row_add)
    # planted mistake: rows inserted before the columns lose the corner the rows own
    scr = big[5:25, 4:28]
    want = big[7:27, 7:31]
    cols, rows = want[:, W - nc:].T.copy(), want[H - nr:].copy()
    cols[:, H - nr:] = 99.0   # the corner as the columns see it differs from the rows'
    out = A.make_layer(scr, cols, rows, -1, -1)
    assert np.array_equal(out, want)
    bad = out.copy()
    bad[:, W - nc:] = cols.T
    assert not np.array_equal(bad, want)


def test_create_pupil_phs_piston_pupil_and_scale():
    rng = np.random.default_rng(5)
    npup = 16
    scr = {"a": rng.standard_normal((24, 24)), "b": rng.standard_normal((22, 26))}
    layers = [dict(key="a", x0=2, y0=3, pos_col=0.0, pos_row=0.0, col_add=-1, row_add=-1, interp_col=0.0,
                   interp_row=0.0),
              dict(key="b", x0=1, y0=2, pos_col=0.3, pos_row=0.6, col_add=1, row_add=-1, interp_col=0.25,
                   interp_row=0.5)]
    yy, xx = np.mgrid[0:npup, 0:npup] - (npup - 1) / 2.
    pup = (np.hypot(yy, xx) <= npup / 2.).astype(np.float32)
    out = A.create_pupil_phs(scr, layers, npup, pupfn=pup, remove_piston=True, source_lam=640.)
    assert abs((out * pup).sum()) < 1e-9 * np.abs(out).sum()
    assert np.all(out[pup == 0] == 0)
    # one layer at whole-pixel positions: the window, piston removed, scaled
    one = [dict(layers[0])]
    got = A.create_pupil_phs(scr, one, npup, pupfn=None, remove_piston=False, source_lam=640.)
    # add at end with interp 0 moves one whole pixel: window starts at (y0 + 1, x0 + 1)
    want = scr["a"][4:4 + npup, 3:3 + npup] * (500. / 640.)
    assert np.allclose(got, want, rtol=0, atol=1e-12)
    # planted mistake: scaling by lambda/500 instead of 500/lambda
    assert not np.allclose(got, scr["a"][4:4 + npup, 3:3 + npup] * (640. / 500.))


def test_natural_spline_against_scipy():
    from scipy.interpolate import CubicSpline
    rng = np.random.default_rng(7)
    x = np.cumsum(rng.uniform(0.5, 2.0, 9))
    y = rng.standard_normal(9)
    xe = np.linspace(x[0], x[-1], 101)
    want = CubicSpline(x, y, bc_type="natural")(xe)
    got = A.natural_spline(x, y, xe)
    assert np.abs(got - want).max() < 1e-12
    # interpolates the knots and reproduces a straight line
    assert np.abs(A.natural_spline(x, y, x) - y).max() < 1e-13
    assert np.abs(A.natural_spline(x, 2 * x - 1, xe) - (2 * xe - 1)).max() < 1e-12
    # planted mistake: a clamped (zero-slope) end instead of the natural end differs
    wrong = CubicSpline(x, y, bc_type="clamped")(xe)
    assert np.abs(wrong - got).max() > 1e-3


def test_cub_spline_interp_separable():
    rng = np.random.default_rng(9)
    yin = np.arange(7.) * 2.5
    xin = np.arange(6.) * 3.0
    fy = rng.standard_normal(7)
    gx = rng.standard_normal(6)
    a = np.outer(fy, gx)
    yout = np.linspace(0, yin[-1], 23)
    xout = np.linspace(0, xin[-1], 19)
    got = A.cub_spline_interp(a, yin, xin, yout, xout, np.float64)
    want = np.outer(A.natural_spline(yin, fy, yout), A.natural_spline(xin, gx, xout))
    assert np.abs(got - want).max() < 1e-12
    # planted mistake: the passes' axes exchanged (y knots used along x) differs
    wrong = np.outer(A.natural_spline(xin, fy[:6], np.linspace(0, xin[-1], 23)), A.natural_spline(yin, np.r_[gx, 0.], np.linspace(0, yin[-1], 19)))
    assert np.abs(wrong - got).max() > 1e-3


def test_fudge_hand_case():
    a = np.arange(9, dtype=np.float32).reshape(3, 3)
    got = A.fudge(a, 0.5)
    # neighbour means: corners /2, edges /3, centre /4
    mean = np.array([[(1 + 3) / 2, (0 + 2 + 4) / 3, (1 + 5) / 2],
                     [(0 + 4 + 6) / 3, (1 + 3 + 5 + 7) / 4, (2 + 4 + 8) / 3],
                     [(3 + 7) / 2, (4 + 6 + 8) / 3, (5 + 7) / 2]], np.float64)
    want = a + (mean - a) * 0.5
    assert np.abs(got - want).max() < 1e-6
    assert got.dtype == np.float32
    c = np.full((5, 5), 2.5, np.float32)
    assert np.array_equal(A.fudge(c, 0.1), c)
    assert A.fudge(a, 0) is a
    # planted mistake: every neighbour sum divided by 4 misses the edges' and corners' counts
    wrong = a + (np.where(True, mean * 0 + (mean * np.array([[2, 3, 2], [3, 4, 3], [2, 3, 2]]) / 4), 0) - a) * 0.5
    assert np.abs(wrong - got).max() > 0.1


def test_mirror_fit_spline_constant_and_knots():
    nact, npup = 9, 33
    act = np.full((nact, nact), 0.7, np.float32)
    surf = A.mirror_fit_spline(act, npup)
    assert np.abs(surf - np.float32(0.7)).max() < 1e-6
    # on a fried geometry with npup - 1 a multiple of nact - 1 the knots are pixels: the surface
    # there is the (coupled) actuator map
    rng = np.random.default_rng(11)
    act = rng.standard_normal((nact, nact)).astype(np.float32)
    surf = A.mirror_fit_spline(act, npup, act_coupling=0.1)
    step = (npup - 1) // (nact - 1)
    assert np.abs(surf[::step, ::step] - A.fudge(act, 0.1)).max() < 1e-6


def test_select_sub_pupil_piston_and_pupil():
    rng = np.random.default_rng(13)
    npup = 16
    yy, xx = np.mgrid[0:npup, 0:npup] - (npup - 1) / 2.
    pup = (np.hypot(yy, xx) <= npup / 2.).astype(np.float32)
    atm = rng.standard_normal((npup, npup)).astype(np.float32)
    dm = rng.standard_normal((npup + 4, npup + 4)).astype(np.float32)
    out = A.select_sub_pupil(atm, dm, 2, 1, pupfn=pup, wavelength_adjustor=0.8, conj_height=0.)
    assert abs(float(out.astype(np.float64).sum())) < 1e-4 * np.abs(out).sum()
    assert np.all(out[pup == 0] == 0)
    raw = (atm + (dm[2:2 + npup, 1:1 + npup] * np.float32(0.8))) * pup
    assert np.allclose(out, (raw - raw.sum(dtype=np.float64) / pup.sum()) * pup, atol=1e-5)
    # a negative conjugate height subtracts the mirror
    neg = A.select_sub_pupil(atm, dm, 2, 1, pupfn=pup, wavelength_adjustor=0.8, conj_height=-1.)
    raw2 = (atm - (dm[2:2 + npup, 1:1 + npup] * np.float32(0.8))) * pup
    assert np.allclose(neg, (raw2 - raw2.sum(dtype=np.float64) / pup.sum()) * pup, atol=1e-5)
    assert not np.allclose(neg, out, atol=1e-3)
