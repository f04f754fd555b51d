"""NumPy oracle for the phase producer upstream of the Shack-Hartmann path: the atmosphere's pupil
phase and the deformable mirror's surface on the line of sight.

TEST INFRASTRUCTURE. This module is the checker for the HIP atmosphere and DM kernels of
``dasp_amd``. Only ``tests/`` may import it; the product never imports it.

Restated from the agb32/dasp source tree (citations relative to it):
  linearshift            cmod/src/interpmodule.c:1234-1300 (interp_linearshift, float64 path)
  layer_window_shift     util/atmos.py:697-725 (createPupilPhs: integer and fractional shifts)
  create_pupil_phs       util/atmos.py:678-833 for a natural guide star (sourceAlt 0, zenith 0, no
                         uplink tip-tilt): layers summed in height order, piston, pupil, 500/lambda
  make_layer             science/infAtmos.py:364-402 (makeLayer: the screen shifted and the new
                         columns then rows inserted)
  natural_spline_*       GSL's gsl_interp_cspline (natural cubic spline) as cmod/src/interpmodule.c:
                         294-600 (gslCubSplineIntrp) uses it: second derivatives from the
                         tridiagonal system with zero end curvature, then the cubic on each interval
  cub_spline_interp      gslCubSplineIntrp: the first pass along y for every input column, the
                         second along x for every output row, double precision, result cast to the
                         output dtype
  fudge                  util/dm.py:1723-1740 (MirrorSurface.fudge, the actuator coupling)
  mirror_fit_spline      util/dm.py:1467-1486 (calcCoords, spline branch), :1520-1536
                         (interpolateSpline)
  select_sub_pupil       util/dm.py:2057-2151 (DMLineOfSight.selectSubPupil, ground-conjugate DM
                         without sub-pixel interpolation or alignment offset)

How each function is pinned (tests/test_oracle_pins_atmos.py): linearshift reproduces bilinear
surfaces a + b x + c y + d x y exactly and is a plain window copy at zero shift; make_layer equals
cutting the same window out of one longer screen; create_pupil_phs leaves no piston inside the
pupil and, for one layer at whole-pixel shifts, equals the scaled window; the natural spline
equals scipy.interpolate.CubicSpline(bc_type="natural") (an independent implementation) and
reproduces straight lines; cub_spline_interp of a separable product f(y) g(x) is the product of
the one-dimensional splines; fudge leaves constant maps unchanged and matches a hand-computed 3x3
case; select_sub_pupil leaves no piston inside the pupil.
"""
import numpy as np

f64 = np.float64
f32 = np.float32


# ---- atmosphere --------------------------------------------------------------------------------

def linearshift_weights(xshift, yshift):
    """The four weights of interp_linearshift (cmod/src/interpmodule.c:1276-1279): the shifts
    arrive as C floats (PyArg "f"), the products are formed in double and stored as floats."""
    x = f64(f32(xshift))
    y = f64(f32(yshift))
    return (f32((1. - x) * (1. - y)), f32((1. - x) * y), f32(x * (1. - y)), f32(x * y))


def linearshift(inarr, xshift, yshift, ny, nx):
    """out[i, j] = in[i, j] s1 + in[i+1, j] s2 + in[i, j+1] s3 + in[i+1, j+1] s4 in float64, added in
    that order (cmod/src/interpmodule.c:1293-1300): xshift weights the next column, yshift the next
    row."""
    s1, s2, s3, s4 = (f64(s) for s in linearshift_weights(xshift, yshift))
    a = np.asarray(inarr, f64)
    out = a[:ny, :nx] * s1
    out = out + a[1:ny + 1, :nx] * s2
    out = out + a[:ny, 1:nx + 1] * s3
    out = out + a[1:ny + 1, 1:nx + 1] * s4
    return out


def layer_window_shift(pos_start, col_add, interp_pos):
    """util/atmos.py:697-720: the static offset due to the source direction plus the layer's
    current sub-pixel position; returns (whole-pixel shift, fractional shift)."""
    p = pos_start
    if col_add >= 0:
        if interp_pos > 0:
            p += 1 - interp_pos
    else:
        if interp_pos > 0:
            p += interp_pos
        else:
            p += 1
    whole = int(np.floor(p))
    return whole, p - whole


def create_pupil_phs(screens, layers, npup, pupfn=None, remove_piston=True, source_lam=500.,
                     intrinsic=None):
    """util/atmos.py:678-833 for a natural guide star. screens: {key: 2-D float64 screen};
    layers: list (in increasing height) of dicts with keys key, x0, y0 (posDict[0], posDict[1]),
    pos_col, pos_row (posDict[4], posDict[5]), col_add, row_add, interp_col, interp_row (the layer's
    interpPosCol/interpPosRow of this step). pupfn None: control["fullPupil"] (pfn = 1, area npup^2)."""
    out = np.zeros((npup, npup), f64) if intrinsic is None else np.array(intrinsic, f64)
    for L in layers:
        sy, fc = layer_window_shift(L["pos_col"], L["col_add"], L["interp_col"])
        sx, fr = layer_window_shift(L["pos_row"], L["row_add"], L["interp_row"])
        scr = screens[L["key"]]
        r0, c0 = L["y0"] + sx, L["x0"] + sy
        win = scr[r0:r0 + npup + 1, c0:c0 + npup + 1]
        out = out + linearshift(win, fc, fr, npup, npup)
    if pupfn is None:
        pfn, area = 1., float(npup * npup)
    else:
        pfn = np.asarray(pupfn)
        area = float(pfn.sum())
    if remove_piston:
        out = out - np.sum(out * pfn) / area
    if pupfn is not None:
        out = out * pfn
    if source_lam != 500.:
        out = out * (500. / source_lam)
    return out


def make_layer(screen, new_cols, new_rows, col_add, row_add):
    """science/infAtmos.py:364-402: the screen moves by (nrows, ncols) away from the side the new
    data comes in on; the new columns (new_cols: (ncols, rows), as colInput holds them) are written
    first, then the new rows (new_rows: (nrows, cols)) over the full width."""
    s = np.array(screen, f64)
    H, W = s.shape
    nc = new_cols.shape[0]
    nr = new_rows.shape[0]
    out = s.copy()
    if col_add < 0:
        if row_add < 0:
            out[:H - nr, :W - nc] = s[nr:, nc:]
            out[:, W - nc:] = new_cols.T
            out[H - nr:] = new_rows
        else:
            out[nr:, :W - nc] = s[:H - nr, nc:]
            out[:, W - nc:] = new_cols.T
            out[:nr] = new_rows
    else:
        if row_add < 0:
            out[:H - nr, nc:] = s[nr:, :W - nc]
            out[:, :nc] = new_cols.T
            out[H - nr:] = new_rows
        else:
            out[nr:, nc:] = s[:H - nr, :W - nc]
            out[:, :nc] = new_cols.T
            out[:nr] = new_rows
    return out


# ---- deformable mirror -----------------------------------------------------------------------

def natural_spline_coeffs(x, y):
    """Second-derivative coefficients c (the curvature halved, GSL's convention) of the natural
    cubic spline through (x, y): c[0] = c[n-1] = 0 and, for 0 < i < n-1,
    h[i-1] c[i-1] + 2 (h[i-1] + h[i]) c[i] + h[i] c[i+1] = 3 (dy[i]/h[i] - dy[i-1]/h[i-1]),
    solved by elimination down the tridiagonal system."""
    x = np.asarray(x, f64)
    y = np.asarray(y, f64)
    n = x.size
    h = np.diff(x)
    dy = np.diff(y)
    c = np.zeros(n, f64)
    if n < 3:
        return c
    m = n - 2
    diag = 2. * (h[:-1] + h[1:])
    off = h[1:-1]
    rhs = 3. * (dy[1:] / h[1:] - dy[:-1] / h[:-1])
    # Thomas algorithm on the symmetric tridiagonal system
    d = diag.copy()
    r = rhs.copy()
    for i in range(1, m):
        w = off[i - 1] / d[i - 1]
        d[i] -= w * off[i - 1]
        r[i] -= w * r[i - 1]
    sol = np.zeros(m, f64)
    sol[m - 1] = r[m - 1] / d[m - 1]
    for i in range(m - 2, -1, -1):
        sol[i] = (r[i] - off[i] * sol[i + 1]) / d[i]
    c[1:-1] = sol
    return c


def natural_spline_eval(x, y, c, xe):
    """Evaluate the natural cubic spline at xe (inside [x[0], x[-1]]): on [x[i], x[i+1]],
    y[i] + dx (b + dx (c[i] + dx d)) with b = dy/h - h (c[i+1] + 2 c[i]) / 3 and
    d = (c[i+1] - c[i]) / (3 h)."""
    x = np.asarray(x, f64)
    y = np.asarray(y, f64)
    xe = np.asarray(xe, f64)
    i = np.clip(np.searchsorted(x, xe, side="right") - 1, 0, x.size - 2)
    h = x[i + 1] - x[i]
    dx = xe - x[i]
    b = (y[i + 1] - y[i]) / h - h * (c[i + 1] + 2. * c[i]) / 3.
    d = (c[i + 1] - c[i]) / (3. * h)
    return y[i] + dx * (b + dx * (c[i] + dx * d))


def natural_spline(x, y, xe):
    return natural_spline_eval(x, y, natural_spline_coeffs(x, y), xe)


def cub_spline_interp(a, yin, xin, yout, xout, out_dtype=f32):
    """gslCubSplineIntrp (cmod/src/interpmodule.c:294-600): input (nInY, nInX) on knots yin, xin;
    pass 1 splines every input column along y at yout, pass 2 every resulting row along x at xout,
    in double; the result is cast to the output dtype."""
    a = np.asarray(a, f64)
    tmp = np.empty((len(yout), a.shape[1]), f64)
    for j in range(a.shape[1]):
        tmp[:, j] = natural_spline(yin, a[:, j], yout)
    out = np.empty((len(yout), len(xout)), f64)
    for i in range(len(yout)):
        out[i] = natural_spline(xin, tmp[i], xout)
    return out.astype(out_dtype)


def spline_coords(npup, nact, actoffset=0.):
    """MirrorSurface.calcCoords, spline branch (util/dm.py:1483-1485): output pixel coordinates
    x and actuator positions x2."""
    spacing = (npup - 1.) / (nact - 1 + actoffset * 2.)
    x = np.arange(npup).astype(f64)
    x2 = (np.arange(nact) * spacing + spacing * actoffset).astype(f64)
    return x, x2


def fudge(actmap, fiddle=0.1):
    """MirrorSurface.fudge (util/dm.py:1723-1740) in the actuator map's dtype: each actuator moves
    a fraction fiddle toward the mean of its 4, 3 (edges) or 2 (corners) neighbours."""
    if fiddle == 0 or fiddle is None:
        return actmap
    a = np.asarray(actmap)
    dt = a.dtype
    d2 = np.zeros(a.shape, dt)
    d2[1:] = a[:-1]
    d2[:-1] += a[1:]
    d2[:, 1:] += a[:, :-1]
    d2[:, :-1] += a[:, 1:]
    d2[1:-1, 1:-1] /= dt.type(4)
    d2[1:-1, ::d2.shape[1] - 1] /= dt.type(3)
    d2[::d2.shape[0] - 1, 1:-1] /= dt.type(3)
    d2[::d2.shape[0] - 1, ::d2.shape[1] - 1] /= dt.type(2)
    res = a + (d2 - a) * dt.type(fiddle)
    return res.astype(dt)


def mirror_fit_spline(actmap, npup, act_coupling=None, actoffset=0.):
    """MirrorSurface.fit for typ "spline" (util/dm.py:1520-1536): the coupled actuator map
    interpolated onto the npup x npup surface (float32, as phsOut is)."""
    a = fudge(actmap, act_coupling)
    x, x2 = spline_coords(npup, a.shape[0], actoffset)
    return cub_spline_interp(a, x2, x2, x, x, f32)


def select_sub_pupil(output, dmphs, yoff, xoff, pupfn=None, add_to_output=True, conj_height=0.,
                     wavelength_adjustor=1., remove_piston=True):
    """DMLineOfSight.selectSubPupil (util/dm.py:2057-2151) for a ground-conjugate DM without
    sub-pixel interpolation or alignment offset, float32 as xinterp_dm's outputData: the DM's
    selected window added (or subtracted for conjHeight < 0), times the wavelength adjustor, then the
    pupil, the piston inside it removed, and the pupil again. The piston sum is accumulated in
    double here (the reference's float32 array sum rounds differently by about 1e-7 of the sum)."""
    out = np.array(output, f32)
    npup = out.shape[0]
    sel = np.asarray(dmphs, f32)[yoff:yoff + npup, xoff:xoff + npup]
    if wavelength_adjustor != 1:
        sel = (sel * f32(wavelength_adjustor)).astype(f32)
    if conj_height < 0:
        sel = -sel
    out = (out + sel) if add_to_output else sel.copy()
    out = out.astype(f32)
    if pupfn is not None:
        pfn = np.asarray(pupfn, f32)
        out = (out * pfn).astype(f32)
        if remove_piston:
            phasesum = out.sum(dtype=f64)
            out = (out - f32(phasesum / f64(pfn.sum()))).astype(f32)
            out = (out * pfn).astype(f32)
    return out
