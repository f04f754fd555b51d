"""Pins of the FP64 oracle against mathematics and independent routes (CPU only).

The reference ships no numeric golden vectors for this path and cannot be built here, so each
oracle function is checked against a closed form, a brute-force computation, or an invariant
chosen so that a dropped term, a wrong sign/index or a transposed operand fails.
"""
import math

import numpy as np
import pytest

from oracle import centoracle as O
from dasp_amd.util.tel import obstructed_pupil, Pupil


def brute_dft2(E, F):
    n = E.shape[0]
    X = np.zeros((F, F), complex)
    for k in range(F):
        for l in range(F):
            acc = 0j
            for i in range(n):
                for j in range(n):
                    acc += E[i, j] * np.exp(-2j * math.pi * (k * i + l * j) / F)
            X[k, l] = acc
    return X


@pytest.mark.parametrize("n,F,Q", [(4, 8, 8), (3, 8, 6), (4, 6, 6), (2, 4, 7), (4, 8, 4)])
def test_hll_vs_brute_dft(n, F, Q):
    rng = np.random.default_rng(1)
    phs = rng.normal(size=(1, n, n)) * 2
    pup = (rng.random((1, n, n)) > 0.2).astype(float)
    hll = O.compute_hll(phs, pup, F, Q)[0]
    t = (math.pi / F * (F + 1.0)) * (np.add.outer(np.arange(n), np.arange(n)) + 1 - n)
    X = brute_dft2(pup[0] * np.exp(1j * (phs[0] - t)), F)
    P = np.abs(X) ** 2
    ref = np.zeros((Q, Q))
    for i in range(Q):
        for j in range(Q):
            if Q <= F:
                o = (F - Q) // 2
                ref[Q - 1 - i, Q - 1 - j] = P[o + i, o + j]
            elif i < F and j < F:
                r = (Q - F) // 2
                ref[F - 1 - i + r, F - 1 - j + r] = P[i, j]
    np.testing.assert_allclose(hll, ref, rtol=1e-10, atol=1e-9)


def test_parseval_total():
    p = obstructed_pupil(8, 8, 0.25)
    pup = p.perSubap(8, vectorAlign=0).reshape(64, 8, 8)
    rng = np.random.default_rng(2)
    phs = rng.normal(size=(64, 8, 8)) * 3
    hll = O.compute_hll(phs, pup, 16, 16)
    np.testing.assert_allclose(hll.sum(axis=(1, 2)), 256 * pup.sum(axis=(1, 2)), rtol=1e-12)


def _params(nsubx, n, F, Q, M, obs=0.0, **kw):
    pupil = obstructed_pupil(nsubx, n, obs) if obs > 0 else Pupil(nsubx * n, nsubx * n / 2., 0, nsubx)
    if obs == 0:
        pupil.fn[:] = True       # full square subapertures for closed forms
    pupfn = pupil.perSubap(nsubx, vectorAlign=0)
    subflag = pupil.getSubapFlag(nsubx, 0.5)
    frac = (pupil.getSubarea(nsubx) / (n * n)).astype(np.float32)
    return O.Params(nsubx, n, F, Q, M, kw.pop("ncen", M), pupfn, subflag, frac, **kw)


def test_flat_phase_centred():
    p = _params(2, 8, 16, 16, 8)
    c, img, _ = O.run_phase(p, np.zeros((16, 16)))
    assert np.abs(c).max() < 1e-12
    np.testing.assert_allclose(img.sum(axis=(1, 2)), 1.0, rtol=1e-12)   # sig*frac, clip==F


def _dirichlet2(th, n):
    s = np.sin(th / 2)
    out = np.where(np.abs(s) < 1e-14, float(n * n), np.sin(n * th / 2) ** 2 / np.where(np.abs(s) < 1e-14, 1, s) ** 2)
    return out


@pytest.mark.parametrize("shift", [1.0, 0.3, -2.5])
def test_ramp_dirichlet_closed_form(shift):
    """A phase ramp a*j along x: the image is R(k)C(l) with Dirichlet kernels, evaluated here
    without any FFT, then binned and centroided by hand."""
    n, F, M = 8, 16, 8
    a = 2 * math.pi * shift / F
    p = _params(1, n, F, F, M)
    jj = np.arange(n)
    phs = np.tile(a * jj, (n, 1))
    c, img, _ = O.run_phase(p, phs)
    tau = math.pi * (F + 1) / F
    k = np.arange(F)
    R = _dirichlet2(-tau - 2 * math.pi * k / F, n)
    C = _dirichlet2(a - tau - 2 * math.pi * k / F, n)
    hll = np.outer(R[::-1], C[::-1])               # 180-degree rotation, Q == F
    b = F // M
    im = hll.reshape(M, b, M, b).sum(axis=(1, 3))
    im *= 1.0 / (n * n * F * F)
    np.testing.assert_allclose(img[0], im, rtol=1e-10, atol=1e-14)
    cx = (im.sum(0) * np.arange(M)).sum() / im.sum() - (M - 1) / 2
    cy = (im.sum(1) * np.arange(M)).sum() / im.sum() - (M - 1) / 2
    np.testing.assert_allclose(c[0], [cx, cy], atol=1e-12)
    assert abs(cy) < 1e-12
    assert np.sign(c[0, 0]) == -np.sign(shift)      # the image is rotated by 180 degrees
    if shift == 1.0:
        assert abs(c[0, 0] - (-0.4684)) < 1e-3


def test_transpose_swaps_xy():
    p = _params(4, 8, 16, 16, 8, obs=0.2)
    rng = np.random.default_rng(3)
    phs = rng.normal(size=(32, 32))
    c1, _, _ = O.run_phase(p, phs)
    c2, _, _ = O.run_phase(p, phs.T)
    c1 = c1.reshape(4, 4, 2)
    c2 = c2.reshape(4, 4, 2)
    np.testing.assert_allclose(c2[:, :, 0], c1[:, :, 1].T, atol=1e-12)
    np.testing.assert_allclose(c2[:, :, 1], c1[:, :, 0].T, atol=1e-12)


def test_orientation_gradient_like_wfsCent():
    """test/standalone/wfsCent.py:98-115: a gradient along rows makes the y slopes vary more
    than the x slopes, and its transpose the opposite."""
    p = _params(10, 8, 16, 16, 8, obs=1 / 7.)
    npup = 80
    flat, _, _ = O.run_phase(p, np.zeros((npup, npup)))
    grad = np.add.outer(np.linspace(-1, 1, npup), np.zeros(npup)) * 20
    c, _, _ = O.run_phase(p, grad)
    d = (c - flat)[p.subflag == 1]
    assert d[:, 0].var() < d[:, 1].var()
    c, _, _ = O.run_phase(p, grad.T)
    d = (c - flat)[p.subflag == 1]
    assert d[:, 0].var() > d[:, 1].var()


def test_prebin_block_sums():
    rng = np.random.default_rng(4)
    h = rng.random((3, 16, 16))
    out = O.prebin_image(h, 16, 2, 8)
    np.testing.assert_allclose(out, h.reshape(3, 8, 2, 8, 2).sum(axis=(2, 4)))
    out = O.prebin_image(h, 16, 2, 12)       # pad 8 -> 12 centred
    assert np.allclose(out[:, 2:10, 2:10], h.reshape(3, 8, 2, 8, 2).sum(axis=(2, 4)))
    assert out[:, :2].sum() == 0 and out[:, 10:].sum() == 0
    out = O.prebin_image(h, 16, 2, 6)        # clip 8 -> 6: the offset is in unbinned pixels
    np.testing.assert_allclose(out, h[:, 1:13, 1:13].reshape(3, 6, 2, 6, 2).sum(axis=(2, 4)))


def _flip_literal(a, mirror):
    """flipQuad / flipArray written as the C loops (swap pairs)."""
    a = a.copy()
    n = a.shape[0]

    def quad(r0, c0, h, w):
        for i in range(h // 2):
            for j in range(w):
                a[r0 + i, c0 + j], a[r0 + h - 1 - i, c0 + w - 1 - j] = a[r0 + h - 1 - i, c0 + w - 1 - j], a[r0 + i, c0 + j]
        if h % 2:
            i = h // 2
            for j in range(w // 2):
                a[r0 + i, c0 + j], a[r0 + h - 1 - i, c0 + w - 1 - j] = a[r0 + h - 1 - i, c0 + w - 1 - j], a[r0 + i, c0 + j]
    w1 = n // 2
    w3 = (n + 1) // 2
    if mirror == 0 or n % 2:
        quad(0, 0, w1, w1)
        quad(w1, 0, w3, w1)
        quad(0, w1, w1, w3)
        quad(w1, w1, w3, w3)
        if mirror:
            quad(0, 0, n, n)
        return a
    f2 = n // 2
    for i in range(f2):
        t = a[i, :f2].copy(); a[i, :f2] = a[i + f2, f2:]; a[i + f2, f2:] = t
        t = a[i, f2:].copy(); a[i, f2:] = a[i + f2, :f2]; a[i + f2, :f2] = t
    return a


@pytest.mark.parametrize("n", [4, 5, 6, 7, 8])
def test_flip_array(n):
    a = np.arange(n * n, dtype=float).reshape(n, n)
    for m in (0, 1):
        np.testing.assert_array_equal(O.flip_array(a, m), _flip_literal(a, m))
    if n % 2 == 0:
        np.testing.assert_array_equal(O.flip_array(a, 1), np.fft.fftshift(a))


def test_convolution_brute():
    P, C = 8, 6
    rng = np.random.default_rng(5)
    h = rng.random((1, P, P))
    spot = rng.random((P, P))
    otf = O.prepare_spot_otf(spot)[None]
    out = O.do_convolution(h, otf, P, P)[0]
    fs = np.fft.fftshift(spot)
    ref = np.zeros((P, P))
    for r in range(P):
        for c in range(P):
            ref[r, c] = sum(h[0, u, v] * fs[(r - u) % P, (c - v) % P] for u in range(P) for v in range(P))
    np.testing.assert_allclose(out, ref * P * P, rtol=1e-10)
    out2 = O.do_convolution(h, otf, P, C)[0]
    np.testing.assert_allclose(out2, (ref * P * P)[1:7, 1:7], rtol=1e-10)
    d = np.zeros((P, P)); d[P // 2, P // 2] = 1
    np.testing.assert_allclose(O.do_convolution(h, O.prepare_spot_otf(d)[None], P, P)[0], h[0] * P * P, rtol=1e-12)


@pytest.mark.parametrize("Q,M", [(16, 8), (32, 16), (64, 8), (12, 5), (16, 6), (10, 7), (8, 8)])
def test_binning(Q, M):
    rng = np.random.default_rng(6)
    x = rng.random((2, Q, Q))
    lit = O.bin_image_literal(x, Q, M)
    A = O.bin_matrix(Q, M)
    np.testing.assert_allclose(np.einsum('ai,sij,bj->sab', A, x, A), lit, rtol=1e-12, atol=1e-14)
    if Q % M == 0:
        b = Q // M
        np.testing.assert_allclose(lit, x.reshape(2, M, b, M, b).sum(axis=(2, 4)), rtol=1e-12)


def test_cog_delta_and_brightest():
    M = 8
    for (r, c) in [(0, 0), (3, 5), (7, 2)]:
        img = np.zeros((M, M)); img[r, c] = 2.0
        assert O.compute_cog(img, M, M) == (c - 3.5, r - 3.5)
        # centre-on-brightest sums absolute indices and then adds the window start again
        # (cmod/src/centmodule.c:941-969): the result is offset by the window start.
        sx = min(max(int(c - 2 + 0.5), 0), M - 4)
        sy = min(max(int(r - 2 + 0.5), 0), M - 4)
        assert O.compute_cog(img, M, 4, None, 1) == pytest.approx((c + sx - 3.5, r + sy - 3.5))
    img = np.zeros((M, M))
    assert O.compute_cog(img, M, M) == (0.0, 0.0)
    sym = np.outer(np.hanning(M), np.hanning(M))
    assert np.allclose(O.compute_cog(sym, M, 6), (0, 0), atol=1e-15)


@pytest.mark.parametrize("nimg,cs", [(8, 8), (6, 10), (16, 16), (5, 7)])
def test_correlation_routes(nimg, cs):
    rng = np.random.default_rng(7)
    img = rng.random((nimg, nimg))
    ref = rng.random((cs, cs))
    hc = O.transform_psf(ref).astype(np.float64)
    got = O.calc_correlation_hc(img, hc, nimg, cs)
    pad = np.zeros((cs, cs)); o = (cs - nimg) // 2; pad[o:o + nimg, o:o + nimg] = img
    np.testing.assert_allclose(got, O.correlate_fft(pad, ref), rtol=2e-6, atol=1e-4)
    fs = np.fft.fftshift(ref)
    brute = np.zeros((cs, cs))
    for r in range(cs):
        for c in range(cs):
            brute[r, c] = sum(pad[u, v] * fs[(u - r) % cs, (v - c) % cs] for u in range(cs) for v in range(cs))
    np.testing.assert_allclose(got, brute * cs * cs, rtol=2e-6, atol=1e-4)
    K = O.hc_pattern_to_spectrum(hc)
    np.testing.assert_allclose(np.fft.ifft2(np.fft.fft2(pad) * K).real * cs * cs, got, rtol=1e-9, atol=1e-9)


def test_identical_correlation_peak_offset():
    """Identical image and reference: the correlation peak sits at N/2 and the fitted centroid
    reads +0.5 before reference subtraction."""
    N = 16
    y, x = np.mgrid[0:N, 0:N]
    img = np.exp(-((x - 7.5) ** 2 + (y - 7.5) ** 2) / 4.0)
    corr = O.calc_correlation_hc(img, O.transform_psf(img).astype(float), N, N)
    assert np.unravel_index(np.argmax(corr), corr.shape) == (N // 2, N // 2)
    cx, cy = O.compute_fit(corr, N, N, 5)
    assert abs(cx - 0.5) < 1e-6 and abs(cy - 0.5) < 1e-6


def test_fit_matrix_and_vertex():
    for fs in (3, 5, 7):
        mx = np.zeros((6, 6))
        for y in range(fs):
            for x in range(fs):
                v = np.array([x * x, x * y, y * y, x, y, 1.0])
                mx += np.outer(v, v)
        np.testing.assert_allclose(O.make_fit_mx(fs), np.linalg.inv(mx)[:5], rtol=1e-9, atol=1e-12)
    N, K = 16, 10
    for (x0, y0) in [(7.3, 8.6), (5.9, 6.2)]:
        y, x = np.mgrid[0:N, 0:N].astype(float)
        img = 100 - (1.3 * (x - x0) ** 2 + 0.4 * (x - x0) * (y - y0) + 0.9 * (y - y0) ** 2)
        cx, cy = O.compute_fit(img, N, K, 5)
        assert abs(cx - (x0 + 0.5 - N / 2)) < 1e-9 and abs(cy - (y0 + 0.5 - N / 2)) < 1e-9


def test_poisson_small_exact_and_large_moments():
    rng = np.random.default_rng(8)
    z = rng.standard_normal(400000)
    for lam in (0.3, 2.0, 9.5):
        k = O.poisson_from_normal(np.full(z.shape, lam), z)
        kk = np.arange(40)
        pmf = np.exp(-lam + kk * math.log(lam) - np.array([math.lgamma(v + 1) for v in kk]))
        hist = np.bincount(k.astype(int), minlength=40)[:40] / z.size
        assert np.abs(hist - pmf).max() < 4 * math.sqrt(pmf.max() / z.size)
    # exactness: the inverse CDF at u = Phi(z) places z on the right side of each CDF step
    lam = 4.0
    zz = np.linspace(-4, 4, 2001)
    k = O.poisson_from_normal(np.full(zz.shape, lam), zz)
    u = O.normal_cdf(zz)
    cdf = np.cumsum([math.exp(-lam) * lam ** j / math.factorial(j) for j in range(40)])
    for kv, uv in zip(k.astype(int), u):
        assert uv <= cdf[kv] + 1e-12 and (kv == 0 or uv > cdf[kv - 1] - 1e-12)
    for lam in (10.0, 50.0, 300.0):
        k = O.poisson_from_normal(np.full(z.shape, lam), z)
        assert abs(k.mean() - lam) < 5 * math.sqrt(lam / z.size) + 0.01
        assert abs(k.var() / lam - 1) < 0.02
        # total-variation distance of the Cornish-Fisher map to the exact pmf
        nq = 1000000
        qs = (np.arange(nq) + 0.5) / nq
        from scipy.special import ndtri
        kq = O.poisson_from_normal(np.full(qs.shape, lam), ndtri(qs)).astype(int)
        hist = np.bincount(kq, minlength=int(lam * 3)) / qs.size
        kk = np.arange(hist.size)
        pmf = np.exp(-lam + kk * math.log(lam) - np.array([math.lgamma(v + 1) for v in kk]))
        # the quantile grid itself contributes about (support width)/nq
        assert 0.5 * np.abs(hist - pmf).sum() < 0.045 / lam + 12 * math.sqrt(lam) / nq


def test_interpolate_phase_reproduces_quadratics():
    m, nsubx, n = 20, 3, 8
    y, x = np.mgrid[0:m, 0:m].astype(float)
    inp = 0.3 * x + 0.7 * y + 0.01 * x * x - 0.02 * x * y + 2
    out = O.interpolate_phase(inp, nsubx, n)
    tot = nsubx * n
    sc = (m - 1.0) / (tot - 1)
    yy, xx = np.mgrid[0:tot, 0:tot].astype(float) * sc
    ref = 0.3 * xx + 0.7 * yy + 0.01 * xx * xx - 0.02 * xx * yy + 2
    np.testing.assert_allclose(out[2:-2, 2:-2], ref[2:-2, 2:-2], atol=1e-9)


def test_brute_and_diff_correlation():
    rng = np.random.default_rng(9)
    N, K = 8, 6
    img = rng.random((N, N))
    b = O.brute_correlation(img, img, N, N, K)
    # zero shift (i=j=K/2) is the full overlap sum of squares
    o = (N - K) // 2
    assert b[K // 2 + o, K // 2 + o] == pytest.approx((img * img).sum())
    d = O.diff_correlation(img, img, N, N, K, 1)
    assert d[K // 2 + o, K // 2 + o] == 0.0
    nz = d[o:o + K, o:o + K]
    assert (nz[nz != 0] > 0).all()


def test_gaussian_fit_vertex_and_in_place_log():
    """computeFit's Gaussian branch fits the parabola to log(img): a sampled 2-D Gaussian is
    recovered exactly; and the transform is applied to the output buffer (cmod/src/centmodule.c
    :1087-1098 on bimg, :1919), so the returned image is the logarithm of the plain one."""
    N, K = 16, 12
    y, x = np.mgrid[0:N, 0:N].astype(float)
    x0, y0 = 7.35, 8.1
    img = 50 * np.exp(-((x - x0) ** 2 / 6.0 + (y - y0) ** 2 / 4.0))
    cx, cy = O.compute_fit(img, N, K, 5, do_gaussian=1, gmin=1e-9, grep=-30.)
    assert abs(cx - (x0 + 0.5 - N / 2)) < 1e-9 and abs(cy - (y0 + 0.5 - N / 2)) < 1e-9
    nsubx, n, nfft, nimg = 2, 8, 16, 8
    pup = obstructed_pupil(nsubx, n, 0.0, 0.5)
    pupfn = pup.perSubap(nsubx, vectorAlign=0)
    flag = pup.getSubapFlag(nsubx, 0.5)
    frac = (pup.getSubarea(nsubx) / (n * n)).astype(np.float32)
    phs = np.random.default_rng(4).normal(size=(nsubx * n, nsubx * n)) * 0.3
    plain = O.Params(nsubx, n, nfft, nfft, nimg, nimg, pupfn, flag, frac, sig=100.0)
    gauss = O.Params(nsubx, n, nfft, nfft, nimg, nimg, pupfn, flag, frac, sig=100.0, gaussianFit=1,
                     gaussianMinVal=1e-3, gaussianReplaceVal=-12.0)
    _, b0, _ = O.run_phase(plain, phs)
    _, b1, _ = O.run_phase(gauss, phs)
    act = flag.ravel() == 1
    want = np.where(b0 < 1e-3, -12.0, np.log(np.maximum(b0, 1e-300)))
    np.testing.assert_allclose(b1[act], want[act], rtol=1e-12, atol=1e-12)


def test_interpolate_phase_edges_and_samples():
    """Edges of interpolatePhase: the output corners are the input corners (the last sample is
    taken at t = 1, :1653-1656), a constant grid stays constant through the replicated edge
    rows/columns, and where the output grid hits input samples exactly (m - 1 a multiple of
    tot - 1) the interpolation returns the samples."""
    rng = np.random.default_rng(9)
    for m, nsubx, n in ((13, 2, 4), (9, 3, 3), (40, 4, 8)):
        inp = rng.normal(size=(m, m))
        out = O.interpolate_phase(inp, nsubx, n)
        for (a, b) in ((0, 0), (0, -1), (-1, 0), (-1, -1)):
            assert abs(out[a, b] - inp[a, b]) < 1e-12
        np.testing.assert_allclose(O.interpolate_phase(np.full((m, m), 2.5), nsubx, n), 2.5, rtol=0, atol=1e-12)
    inp = rng.normal(size=(25, 25))
    out = O.interpolate_phase(inp, 1, 13)   # tot = 13, scale = 2: output k samples input 2k
    np.testing.assert_allclose(out, inp[::2, ::2], rtol=0, atol=1e-12)
