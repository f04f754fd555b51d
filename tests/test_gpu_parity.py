"""GPU parity: the gfx950 engine, called through the drop-in centcmod contract, against the FP64
oracle on the same phase and the same pre-drawn noise.

Bar: noiseless centroids within 1e-4 px absolute (FP32 engine vs FP64 oracle). With thresholds
or noise, a pixel can sit on a decision boundary (floor, Poisson CDF step) where FP32 and FP64
round differently; those tests bound the number of differing pixels and check the rest to 1e-4.
"""
import numpy as np
import pytest

from dasp_amd import workloads as W
from oracle import centoracle as O
from tests._cases import Case, report, say

pytestmark = pytest.mark.gpu


def _phase(w, frame=0):
    return W.WFS.phase(w, frame)


def _active(case):
    return case.cc.subflag.ravel() == 1


PATHS = {"fused": 0, "general": 3, "fused_fasttrig": 10}


@pytest.mark.parametrize("path", sorted(PATHS))
@pytest.mark.parametrize("wfs", [W.SCAO8, W.WFS80], ids=["scao8x8", "wfs80x80"])
def test_noiseless_cog_vs_oracle(wfs, path):
    say("case noiseless CoG %s %s" % (wfs.name, path))
    c = Case(wfs.nsubx, wfs.n, wfs.nfft, wfs.nimg, obs=wfs.obs, phs=wfs.phase(), form=PATHS[path])
    oc, ob, _ = c.run_both()
    act = _active(c)
    d = report(wfs.name, c.cents, oc, act)
    assert d.max() < 1e-4
    assert np.all(c.cents.reshape(-1, 2)[~act] == 0)
    b = c.bimg.reshape(-1, wfs.nimg, wfs.nimg)[act]
    np.testing.assert_allclose(b, ob[act], rtol=0, atol=2e-5 * ob[act].max())
    np.testing.assert_allclose(b.sum(axis=(1, 2)), ob[act].sum(axis=(1, 2)), rtol=2e-5)
    c.free()


def test_flat_phase_reference_and_refcents():
    say("case flat phase + REFCENTS")
    w = W.SCAO8
    c = Case(w.nsubx, w.n, w.nfft, w.nimg, obs=w.obs)
    c.run_both()
    ref = c.cents.copy()
    assert np.abs(ref).max() < 1e-5          # symmetric pupil, centred spot
    c.phs[:] = w.phase()
    c.cc.update(18, ref)                     # REFCENTS
    c.params.refCents = ref
    oc, _, _ = c.run_both()
    assert report("refcents", c.cents, oc, _active(c)).max() < 1e-4
    c.free()


@pytest.mark.parametrize("path", sorted(PATHS))
def test_noise_poisson_read_relative_floor(path):
    """40x40 WFS, 1000 photons, Poisson + read noise (sigma 0.5) and a 0.1*max floor."""
    w = W.NOISE40
    say("case noise %s %s" % (w.name, path))
    c = Case(w.nsubx, w.n, w.nfft, w.nimg, obs=w.obs, sig=w.sig, readnoise=w.readnoise, addPoisson=1,
             noiseFloor=w.noiseFloor, threshType=0, relativeFloor=True, phs=w.phase(), seed=5, form=PATHS[path])
    oc, ob, _ = c.run_both()
    _check_noisy(c, oc, ob)
    c.free()


def _check_noisy(c, oc, ob):
    """Noisy images against the oracle: a few pixels may round to the neighbouring count (FP32
    against FP64 at a step of the inverse CDF); the other sub-apertures must match."""
    act = _active(c)
    b = c.bimg.reshape(-1, c.nimg, c.nimg)[act]
    diffpix = np.abs(b - ob[act]) > 1e-3 * np.maximum(1, np.abs(ob[act]))
    say("noisy pixels differing: %d of %d" % (diffpix.sum(), diffpix.size))
    assert diffpix.sum() <= max(3, 1e-4 * diffpix.size)
    clean = ~diffpix.reshape(diffpix.shape[0], -1).any(axis=1)
    d = np.abs(c.cents.reshape(-1, 2)[act] - oc[act])
    say("subaps without a flipped pixel: %d of %d; their max|d|=%.3g px" % (clean.sum(), act.sum(), d[clean].max()))
    assert d[clean].max() < 1e-4
    # the noisy images are integers plus read noise: check the photon count statistics
    assert abs(b.sum() / act.sum() - ob[act].sum() / act.sum()) < 5


FUSED_SHAPES = {
    # name: (nsubx, n, nfft, nimg), fused kernel <NP, F, F/nimg>; detector pixels per lane
    "n8_f16_b1": (8, 8, 16, 16),      # 32 per lane (more than one subaperture slot of LDS)
    "n16_f32_b1": (6, 16, 32, 32),    # 64 per lane (more than one subaperture slot of LDS)
    "n8_f16_b2": (8, 8, 16, 8),       # 8 per lane
    "n16_f32_b4": (6, 16, 32, 8),     # 8 per lane, two lanes per detector row
    "n32_f64_b4": (6, 32, 64, 16),    # 16 per lane
    "n32_f64_b2": (4, 32, 64, 32),    # 32 per lane
}


@pytest.mark.parametrize("kind", ["noise", "skyarr"])
@pytest.mark.parametrize("name", sorted(FUSED_SHAPES))
def test_general_detector_every_fused_shape(name, kind):
    """The fused kernels' general detector on every shape: it stages all the lanes' pixels in the
    wave's LDS region, up to 64 per lane when nimg == fftsize (more than a subaperture's transpose
    slot holds). Poisson + read noise with a relative floor (against the oracle with the flipped-
    pixel bounds), or a sky-background array without noise (exact)."""
    nsubx, n, nfft, nimg = FUSED_SHAPES[name]
    say("case general detector %s %s" % (name, kind))
    w = W.WFS(name, nsubx, n, nfft, nimg, 4.2, 0.2, 0.15, 25.0)
    if kind == "noise":
        c = Case(nsubx, n, nfft, nimg, obs=0.2, sig=500.0, readnoise=0.5, addPoisson=1, noiseFloor=0.05,
                 threshType=0, relativeFloor=True, phs=w.phase(), seed=3)
        oc, ob, _ = c.run_both()
        _check_noisy(c, oc, ob)
    else:
        sky = (np.random.default_rng(2).random((nsubx * nsubx, nimg, nimg)) * 0.5).astype(np.float32)
        c = Case(nsubx, n, nfft, nimg, obs=0.2, sig=200.0, skybrightness=sky, phs=w.phase())
        oc, ob, _ = c.run_both()
        act = _active(c)
        b = c.bimg.reshape(-1, nimg, nimg)[act]
        np.testing.assert_allclose(b, ob[act], rtol=0, atol=2e-5 * ob[act].max())
        assert report("skyarr " + name, c.cents, oc, act).max() < 1e-4
    c.free()


def _brightest_counts(S, M2, seed):
    """Per-sub-aperture counts of every kind: both signs, none (0), the whole image (M2: no
    selection), one pixel, all but one, random in between."""
    ub = np.random.default_rng(seed).integers(-(M2 // 2), M2 // 2, S)
    ub[:6] = [M2, -(M2 - 1), 1, -1, 0, M2 - 1]
    return ub.astype(np.int32)


def _brightest_pair(nsubx, n, nfft, nimg, ub, **kw):
    """The same images without (a count of nimg^2 everywhere: the same detector code, nothing
    selected) and with the selection; returns (case with selection, oracle slopes, oracle images,
    the engine's images before selection)."""
    S, M2 = nsubx * nsubx, nimg * nimg
    c0 = Case(nsubx, n, nfft, nimg, useBrightest=np.full(S, M2, np.int32), form=0, **kw)
    c0.run_both()
    b0 = c0.bimg.reshape(S, nimg, nimg).copy()
    c0.free()
    c = Case(nsubx, n, nfft, nimg, useBrightest=ub, form=0, **kw)
    oc, ob, _ = c.run_both()
    return c, oc, ob, b0


def _check_selection(c, b0, ub):
    """The engine's selected images equal selectBrightestPixels of its own images before
    selection, bit for bit (FP32 subtraction of the next value)."""
    S = c.nsubx * c.nsubx
    ubs = np.broadcast_to(np.asarray(ub), (S,))
    act = _active(c)
    want = np.stack([O.select_brightest(b0[s], int(ubs[s])).astype(np.float32) for s in range(S)])
    got = c.bimg.reshape(S, c.nimg, c.nimg)
    bad = (got != want).any(axis=(1, 2)) & act
    say("selection differing from the oracle's on the engine's own images: %d sub-apertures" % bad.sum())
    assert not bad.any()
    cut_ties = 0
    for s in np.nonzero(act)[0]:
        u = abs(int(ubs[s]))
        x = np.sort(b0[s].ravel())[::-1]
        if 0 < u < x.size and x[u - 1] == x[u]:
            cut_ties += 1
    say("sub-apertures with a tie at the cut: %d" % cut_ties)
    return cut_ties


@pytest.mark.parametrize("name", sorted(FUSED_SHAPES))
def test_brightest_every_fused_shape(name):
    """selectBrightestPixels (cmod/src/centmodule.c:1544-1581) in the fused kernels' general
    detector on every shape (keys in registers up to 16 pixels a lane, in LDS above), with the
    per-sub-aperture count array and readCCD's type-1 floor before it (ties at zero wherever a
    count reaches into the zeroed pixels). The selection must equal the oracle's on the engine's
    own images before selection, bit for bit; against the FP64 oracle end to end, the
    sub-apertures whose cut sits between two values closer than the FP32 image error are left out."""
    nsubx, n, nfft, nimg = FUSED_SHAPES[name]
    say("case brightest %s" % name)
    S, M2 = nsubx * nsubx, nimg * nimg
    w = W.WFS(name, nsubx, n, nfft, nimg, 4.2, 0.2, 0.15, 25.0)
    ub = _brightest_counts(S, M2, 4)
    c, oc, ob, b0 = _brightest_pair(nsubx, n, nfft, nimg, ub, obs=0.2, sig=200.0, noiseFloor=0.05 * 200.0 / M2,
                                    threshType=1, phs=w.phase())
    _check_selection(c, b0, ub)
    import copy
    p0 = copy.copy(c.params)
    p0.useBrightest, p0.useBrightestArr = 0, None
    _, ob0, _ = O.run_phase(p0, c.phs)
    act = _active(c)
    err = np.abs(b0 - ob0)[act].max()            # the FP32 image error before selection
    fragile = np.zeros(S, bool)
    for s in range(S):
        u = abs(int(ub[s]))
        x = np.sort(ob0[s].ravel())[::-1]
        if 0 < u < x.size:
            fragile[s] = 0 < x[u - 1] - x[u] <= 4 * err
    keep = act & ~fragile
    say("FP32 image error %.3g; sub-apertures with the cut inside it: %d of %d" % (err, (act & fragile).sum(),
                                                                                  act.sum()))
    b = c.bimg.reshape(-1, nimg, nimg)
    np.testing.assert_allclose(b[keep], ob[keep], rtol=0, atol=max(4 * err, 1e-6))
    assert report("brightest " + name, c.cents, oc, keep).max() < 1e-4
    c.free()


@pytest.mark.parametrize("ub", [10, -10, 60, -60])
def test_brightest_noisy_ties(ub):
    """Brightest pixels of Poisson images (integer counts: many equal values at and around the
    cut, and below it for the subtracted one) with one count for every sub-aperture: bit for bit
    the oracle's selection on the engine's images before selection; end to end against the FP64
    oracle, where a Poisson count that rounds to the neighbouring step (FP32 against FP64) moves
    the cut of its whole sub-aperture, at most a few such sub-apertures and the others exact."""
    nsubx, n, nfft, nimg = 8, 8, 16, 8
    say("case brightest noisy %d" % ub)
    w = W.WFS("bn", nsubx, n, nfft, nimg, 4.2, 0.2, 0.15, 25.0)
    c, oc, ob, b0 = _brightest_pair(nsubx, n, nfft, nimg, ub, obs=0.2, sig=150.0, addPoisson=1, phs=w.phase(),
                                    seed=11)
    assert _check_selection(c, b0, ub) > 0
    act = _active(c)
    b = c.bimg.reshape(-1, nimg, nimg)[act]
    bad = (np.abs(b - ob[act]) > 1e-3 * np.maximum(1, np.abs(ob[act]))).any(axis=(1, 2))
    say("sub-apertures differing from the FP64 oracle: %d of %d" % (bad.sum(), act.sum()))
    assert bad.sum() <= 2
    d = np.abs(c.cents.reshape(-1, 2)[act] - oc[act])
    assert d[~bad].max() < 1e-4
    c.free()


def test_correlation_parabolic_fit():
    """Nyquist-sampled n=32/nfft=64/nimg=16 WFS, FFT correlation with the flat-phase image as
    the reference (util/centroid.py:1580-1626), 5x5 parabolic fit."""
    say("case correlation + parabolic fit")
    nsubx = 12
    wgeo = W.WFS("corr", nsubx, 32, 64, 16, 8.0, 0.2, 0.15, 25.0)
    ref = W.reference_images(wgeo)
    c = Case(nsubx, 32, 64, 16, obs=0.2, correlationCentroiding=1, corrPattern=ref, parabolicFit=5,
             phs=wgeo.phase())
    oc, ob, ocorr = c.run_both()
    act = _active(c)
    d = report("corr+fit", c.cents, oc, act)
    assert d.max() < 1e-4
    ci = c.corrimg.reshape(-1, 16, 16)[act]
    np.testing.assert_allclose(ci, ocorr[act], rtol=0, atol=1e-5 * np.abs(ocorr[act]).max())
    c.free()


def test_general_path_largest_geometry():
    """The general path at the largest supported sides: nimg 64 with an FFT correlation of 64 x 64
    (the per-item LDS region is sized from the geometry; this one needs more than 64 KB, the
    default limit of dynamically sized LDS), against the FP64 oracle."""
    say("case general path nimg 64, corrsize 64")
    nsubx, n, nfft, nimg = 2, 32, 64, 64
    wgeo = W.WFS("big", nsubx, n, nfft, nimg, 2.0, 0.0, 0.2, 25.0)
    ref = W.reference_images(wgeo)
    c = Case(nsubx, n, nfft, nimg, obs=0.0, correlationCentroiding=1, corrPattern=ref, parabolicFit=5,
             phs=wgeo.phase() * 0.5, form=3)
    oc, ob, ocorr = c.run_both()
    act = _active(c)
    assert act.sum() >= 1
    d = report("general nimg 64", c.cents, oc, act)
    assert d.max() < 1e-4
    b = c.bimg.reshape(-1, nimg, nimg)[act]
    np.testing.assert_allclose(b, ob[act], rtol=0, atol=2e-5 * ob[act].max())
    ci = c.corrimg.reshape(-1, nimg, nimg)[act]
    np.testing.assert_allclose(ci, ocorr[act], rtol=0, atol=1e-5 * np.abs(ocorr[act]).max())
    c.free()


@pytest.mark.parametrize("fit", [0, 3], ids=["centred_cog", "parabolic3"])
@pytest.mark.parametrize("nsubx", [5, 7])
def test_correlation_padded_odd_count(nsubx, fit):
    """FFT correlation with a 16x16 pattern on 8x8 images (centre padding, calcCorrelation
    cmod/src/centmodule.c:1403-1486) and sub-aperture counts that leave an unpaired image at the
    end of a slice: the images are correlated two per lane group, the last one alone."""
    say("case correlation padded nsubx=%d fit=%d" % (nsubx, fit))
    nimg, C = 8, 16
    wgeo = W.WFS("corrpad", nsubx, 8, 16, nimg, 4.2, 0.25, 0.2, 30.0)
    ref = W.reference_images(wgeo)
    o = (C - nimg) // 2
    refp = np.zeros((nsubx, nsubx, C, C), np.float32)
    refp[:, :, o:o + nimg, o:o + nimg] = ref
    c = Case(nsubx, 8, 16, nimg, obs=0.25, correlationCentroiding=1, corrPattern=refp, parabolicFit=fit,
             phs=wgeo.phase() * 0.5)
    oc, ob, ocorr = c.run_both()
    act = _active(c)
    d = report("corr padded", c.cents, oc, act)
    assert d.max() < 1e-4
    assert np.all(c.cents.reshape(-1, 2)[~act] == 0)
    ci = c.corrimg.reshape(-1, C, C)[act]
    np.testing.assert_allclose(ci, ocorr[act], rtol=0, atol=1e-5 * np.abs(ocorr[act]).max())
    c.free()


def test_fit_window_outside_ncen_refused():
    """computeFit clamps its fitsize window to the central ncen x ncen pixels, which only works when
    the window fits there (cmod/src/centmodule.c:1069-1086; the reference reads past the image
    otherwise): a 5x5 fit with ncen 4 is refused when the tables are built, before any kernel."""
    from dasp_amd.cmod import cent
    say("case fit window larger than ncen")
    w = W.SCAO8
    with pytest.raises((cent.error, RuntimeError)):
        Case(w.nsubx, w.n, w.nfft, w.nimg, obs=w.obs, ncen=4, parabolicFit=5, phs=w.phase())
    c = Case(w.nsubx, w.n, w.nfft, w.nimg, obs=w.obs, ncen=5, parabolicFit=5, phs=w.phase())
    oc, _, _ = c.run_both()
    assert report("5x5 fit in ncen 5", c.cents, oc, _active(c)).max() < 1e-4
    c.free()


LGS_CASES = {
    # name: (nsubx, n, nfft, nimg, spot size, spot layout, extra)
    "n16_p32_4d": (10, 16, 32, 16, 32, 4, {}),              # fused (psfsize == clipsize == fftsize)
    "n16_p32_2d": (10, 16, 32, 16, 32, 2, {}),
    "n8_p16_4d": (8, 8, 16, 8, 16, 4, {}),
    "n16_p32_noise": (10, 16, 32, 16, 32, 4, dict(readnoise=0.5, addPoisson=1, noiseFloor=0.05, threshType=0, relativeFloor=True)),
    "n16_p64_4d": (6, 16, 32, 16, 64, 4, {}),              # psfsize > fftsize: general path only
}


@pytest.mark.parametrize("path", sorted(PATHS))
@pytest.mark.parametrize("name", sorted(LGS_CASES))
def test_lgs_spot_convolution(name, path):
    say("case LGS spot convolution %s %s" % (name, path))
    nsubx, n, nfft, nimg, P, layout, kw = LGS_CASES[name]
    wgeo = W.WFS("lgs", nsubx, n, nfft, nimg, 8.0, 0.2, 0.15, 25.0, spot=True, spot_size=P, sig=1000.0)
    spots = wgeo.spots()
    if layout == 2:
        spots = np.ascontiguousarray(spots[nsubx // 2, nsubx // 3])
    c = Case(nsubx, n, nfft, nimg, obs=0.2, sig=1000.0, spotpsf=spots, phs=wgeo.phase(), form=PATHS[path],
             seed=7, **kw)
    oc, ob, _ = c.run_both()
    act = _active(c)
    b = c.bimg.reshape(-1, nimg, nimg)[act]
    if kw:
        diffpix = np.abs(b - ob[act]) > 1e-3 * np.maximum(1, np.abs(ob[act]))
        say("noisy pixels differing: %d of %d" % (diffpix.sum(), diffpix.size))
        assert diffpix.sum() <= max(3, 1e-4 * diffpix.size)
        clean = ~diffpix.reshape(diffpix.shape[0], -1).any(axis=1)
        d = np.abs(c.cents.reshape(-1, 2)[act] - oc[act])
        assert d[clean].max() < 1e-4
    else:
        np.testing.assert_allclose(b, ob[act], rtol=0, atol=2e-5 * np.abs(ob[act]).max())
        d = report("lgs " + name, c.cents, oc, act)
        assert d.max() < 1e-4
    c.free()


OPTION_CASES = {
    "pxlpower2": dict(pxlPower=2.),
    "pxlpower1.5": dict(pxlPower=1.5),
    "brightest12": dict(useBrightest=12),
    "brightest-12": dict(useBrightest=-12),
    "sky_scalar_floor0": dict(skybrightness=0.01, noiseFloor=0.02, threshType=0),
    "floor_type1": dict(noiseFloor=0.01, threshType=1),
    "ncen6": dict(ncen=6),
    "clip_fractional_bin": dict(clip=14, nimg=6),
    "optical_binning": dict(opticalBinning=1),
    "imageOnly": dict(imageOnly=1),
    "sigarr": "sigarr",
    "skyarr": "skyarr",
    "centweight4d": "centweight",
    "brute_corr5": "brute",
    "diffcorr2": "diff2",
    "diffcorr4": "diff4",
    "gaussfit": "gauss",
    "quadinterp": "quad",
    "corr1_cog": "corrcog",
    "corr1_fit3_thresh": "corrthresh",
    "parabolic5_image": dict(parabolicFit=5),
    "gauss_image": dict(gaussianFitVals=(1e-4, -12.)),
}


@pytest.mark.parametrize("path", sorted(PATHS))
@pytest.mark.parametrize("name", sorted(OPTION_CASES))
def test_options(name, path):
    say("case option %s %s" % (name, path))
    nsubx, n, nfft, nimg = 6, 8, 16, 8
    wgeo = W.WFS("opt", nsubx, n, nfft, nimg, 4.2, 0.25, 0.2, 30.0)
    phs = wgeo.phase() * 0.5
    S = nsubx * nsubx
    kw = OPTION_CASES[name]
    rng = np.random.default_rng(3)
    if kw == "sigarr":
        kw = dict(sig=(rng.random(S) * 2 + 0.5).astype(np.float32))
    elif kw == "skyarr":
        kw = dict(skybrightness=(rng.random((S, nimg, nimg)) * 0.01).astype(np.float32))
    elif kw == "centweight":
        kw = dict(centWeight=(rng.random((nsubx, nsubx, nimg, nimg)) + 0.5).astype(np.float32))
    elif kw in ("brute", "diff2", "diff4", "gauss", "quad", "corrcog", "corrthresh"):
        ref = W.reference_images(wgeo)
        mode = {"brute": 5, "diff2": 2, "diff4": 4, "gauss": 1, "quad": 1, "corrcog": 1, "corrthresh": 1}[kw]
        kw = dict(correlationCentroiding=mode, corrPattern=ref.astype(np.float32))
        if name == "gaussfit":
            kw.update(gaussianFitVals=(1e-3, -10.))
        if name == "quadinterp":
            kw.update(parabolicFit=-1)
        if name == "corr1_fit3_thresh":
            kw.update(parabolicFit=3, corrThresh=0.3)
    kw = dict(kw)
    nimg_ = kw.pop("nimg", nimg)
    c = Case(nsubx, n, nfft, nimg_, obs=0.25, phs=phs, form=PATHS[path], **kw)
    oc, ob, ocorr = c.run_both()
    act = _active(c)
    b = c.bimg.reshape(-1, nimg_, nimg_)[act]
    np.testing.assert_allclose(b, ob[act], rtol=0, atol=1e-4 * max(1e-30, np.abs(ob[act]).max()))
    if not c.params.imageOnly:
        d = report(name, c.cents, oc, act)
        assert d.max() < 1e-4
    if ocorr is not None and c.corrimg is not None:
        C = c.corrimg.shape[-1]
        ci = c.corrimg.reshape(-1, C, C)[act]
        np.testing.assert_allclose(ci, ocorr[act], rtol=0, atol=1e-4 * max(1e-30, np.abs(ocorr[act]).max()))
    c.free()


def test_integration_cycle():
    """nintegrations=2, nlatency=1: a readout every third call (cmod/src/centmodule.c:1774-1788)."""
    say("case integration cycle")
    w = W.SCAO8
    c = Case(w.nsubx, w.n, w.nfft, w.nimg, obs=w.obs, nintegrations=2, nlatency=1, phs=w.phase())
    outs = []
    for k in range(6):
        c.cents[:] = -7
        c.cc.run(0)
        outs.append(c.cents.copy())
    ro = [np.all(o != -7) for o in outs]
    assert ro == [False, False, True, False, False, True]
    c.free()


def test_image_input_runslope():
    say("case runSlope image input")
    w = W.SCAO8
    src = Case(w.nsubx, w.n, w.nfft, w.nimg, obs=w.obs, phs=w.phase(), sig=100.)
    src.cc.run(0)
    img4 = src.bimg.copy()
    img2 = np.ascontiguousarray(img4.transpose(0, 2, 1, 3).reshape(w.nsubx * w.nimg, w.nsubx * w.nimg))
    src.free()
    for inp in (img2, img4):
        c = Case(w.nsubx, w.n, w.nfft, w.nimg, obs=w.obs, noiseFloor=0.05, inputImage=inp)
        before = c.bimg.copy()
        c.cc.runSlope(0)
        oc, ob, _ = O.run_image(c.params, inp, bimg_in=before)
        assert report("runSlope%dd" % inp.ndim, c.cents, oc, _active(c)).max() < 1e-4
        c.free()


def test_run_device_frames_match_legacy():
    """Device-resident multi-frame step equals per-frame legacy runs."""
    import torch
    from dasp_amd.cmod import cent
    say("case run_device frames")
    w = W.SCAO8
    c = Case(w.nsubx, w.n, w.nfft, w.nimg, obs=w.obs)
    frames = np.stack([w.phase(f) for f in range(3)])
    want = []
    for f in range(3):
        c.phs[:] = frames[f]
        c.cc.run(0)
        want.append(c.cents.copy())
    dev = torch.device("cuda:0")
    ph = torch.from_numpy(frames).to(dev)
    S = w.nsubx * w.nsubx
    ce = torch.zeros((3, S, 2), dtype=torch.float32, device=dev)
    bi = torch.zeros((3, S, w.nimg, w.nimg), dtype=torch.float32, device=dev)
    cent.runDevice(c.cc.centstruct, ph.data_ptr(), w.npup * w.npup, 3, ce.data_ptr(), bi.data_ptr(),
                   stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    got = ce.cpu().numpy().reshape(3, w.nsubx, w.nsubx, 2)
    np.testing.assert_array_equal(got, np.stack(want))
    c.free()


# form codes of tests/_cases.Case: FORM_DFT_FP16 = 2, FORM_DFT_BF16 = 1, the single-product FP16
# form FORM_DFT_FP16_1 = 4; + 10 selects the reduced hardware sin/cos (the library's default trig)
DFT_FORMS = {"fp16": 2, "bf16": 1, "fp16_fasttrig": 12, "bf16_fasttrig": 11, "fp16x1": 4, "fp16x1_fasttrig": 14}


def _strict(form):
    """The compensated FP16 form is held to the FP32 path's 1e-4 px."""
    return form.startswith("fp16") and "x1" not in form


def _img_tol(form):
    """Image tolerance (fraction of the image maximum) of each form: the compensated forms' split
    errors, or one rounding of both operands to FP16 (the single-product form, measured on the
    80x80 nfft32 phase with an FP64 model of the products: 1.3e-3 of the maximum at the median
    sub-aperture, 1.6e-3 at the worst; most of it a common scale, as round-toward-zero shrinks
    every operand)."""
    if "x1" in form:
        return 3e-3
    return 5e-5 if form.startswith("fp16") else 5e-4


def _max_bound(form):
    """Bound on the largest centroid error of the forms not held to 1e-4 px."""
    return 1e-3 if "x1" in form else 1e-2


@pytest.mark.parametrize("form", sorted(DFT_FORMS))
def test_dft_mfma_form_error_vs_fp64(form):
    """The DFT-GEMM form (hi/lo split, FP32 accumulate) on the 80x80 nfft=32 WFS: RMS centroid
    error against FP64 within 1e-3 px (FP16 split is expected near the FP32 path)."""
    w = W.WFS80
    say("case DFT form %s" % form)
    c = Case(w.nsubx, w.n, w.nfft, w.nimg, obs=w.obs, phs=w.phase(), form=DFT_FORMS[form])
    oc, ob, _ = c.run_both()
    act = _active(c)
    d = report("dft-" + form, c.cents, oc, act)
    rms = np.sqrt((d ** 2).mean())
    assert rms < 1e-3
    assert d.max() < (1e-4 if _strict(form) else _max_bound(form))
    c.free()


DFT16_CASES = {
    # name: Case keyword arguments (n=16, nfft=32, nimg=16); the pipelined kernel runs the
    # inline detector, the general one noise and relative floors
    "floor_type0": dict(noiseFloor=0.004, threshType=0),
    "floor_type1": dict(noiseFloor=0.004, threshType=1),
    "ncen12_sky": dict(ncen=12, skybrightness=0.001),
    "noise_relfloor": dict(sig=1000.0, readnoise=0.5, addPoisson=1, noiseFloor=0.1, threshType=0, relativeFloor=True, seed=4),
}


@pytest.mark.parametrize("form", sorted(DFT_FORMS))
@pytest.mark.parametrize("name", sorted(DFT16_CASES))
def test_dft16_options_vs_fp64(name, form):
    """The n=16/nfft=32 DFT-GEMM form with detector options, 11x11 WFS (an odd active count, so
    that XCD slices end in half-empty pairs), images and slopes against the FP64 oracle."""
    say("case DFT16 %s %s" % (name, form))
    wgeo = W.WFS("dft16", 11, 16, 32, 16, 8.0, 0.2, 0.15, 25.0)
    kw = dict(DFT16_CASES[name])
    c = Case(11, 16, 32, 16, obs=0.2, phs=wgeo.phase(), form=DFT_FORMS[form], **kw)
    oc, ob, _ = c.run_both()
    act = _active(c)
    say("active sub-apertures: %d" % act.sum())
    b = c.bimg.reshape(-1, 16, 16)[act]
    tol = _img_tol(form)
    # a pixel within the form's error of a threshold may fall on the other side of it
    diff = np.abs(b - ob[act])
    diffpix = diff > tol * np.abs(ob[act]).max()
    say("pixels beyond %.0e of the maximum: %d of %d" % (tol, diffpix.sum(), diffpix.size))
    if "noise" in name and "x1" in form:
        # one rounding to 16 bits moves the photon rate by up to _img_tol of the maximum, so more
        # pixels draw a neighbouring count from the same normal (as in the n32 noisy case): bound
        # their share and their size (a photon or two, plus the relative floor's shift)
        assert diffpix.sum() <= 2e-2 * diffpix.size
        if diffpix.any():
            say("largest difference of those pixels: %.3g" % diff[diffpix].max())
            assert diff[diffpix].max() <= 2.5
    else:
        assert diffpix.sum() <= max(2, 2e-3 * diffpix.size)
    clean = ~diffpix.reshape(diffpix.shape[0], -1).any(axis=1)
    d = np.abs(c.cents.reshape(-1, 2)[act] - oc[act])[clean]
    assert np.sqrt((d ** 2).mean()) < 1e-3
    if _strict(form) and "noise" not in name:
        assert d.max() < 1e-4
    assert np.all(c.cents.reshape(-1, 2)[~act] == 0)
    c.free()


def test_dft16_refcents():
    """Reference slopes subtracted in the pipelined DFT kernel (REFCENTS update)."""
    from dasp_amd.cmod import cent
    say("case DFT16 refcents")
    wgeo = W.WFS("dft16", 11, 16, 32, 16, 8.0, 0.2, 0.15, 25.0)
    c = Case(11, 16, 32, 16, obs=0.2, form=cent.FORM_DFT_FP16)
    c.run_both()
    ref = c.cents.copy()
    c.phs[:] = wgeo.phase()
    c.cc.update(18, ref)   # REFCENTS
    c.params.refCents = ref
    oc, _, _ = c.run_both()
    assert report("dft16-refcents", c.cents, oc, _active(c)).max() < 1e-4
    c.free()


DFT32_CASES = {
    # name: (nsubx, n, correlation + parabolic fit, noise)
    "n32_cog": (16, 32, False, False),
    "n24_cog": (16, 24, False, False),
    "n32_corr_fit5": (12, 32, True, False),
    "n32_noise_thresh": (12, 32, False, True),
}


@pytest.mark.parametrize("form", sorted(DFT_FORMS))
@pytest.mark.parametrize("name", sorted(DFT32_CASES))
def test_dft32_mfma_form_error_vs_fp64(name, form):
    """The DFT-GEMM form for phase size n <= 32, fftsize 64, nimg 16 (4x4 binning) against FP64:
    noiseless centroids to 1e-4 px (FP16 split) / RMS 1e-3 px (BF16 split); the images the
    correlation and fit stage reads within 5e-5 / 5e-4 of the image maximum."""
    nsubx, n, corr, noise = DFT32_CASES[name]
    say("case DFT32 %s %s" % (name, form))
    wgeo = W.WFS("dft32", nsubx, n, 64, 16, 8.0, 0.2, 0.15, 25.0)
    kw = dict(obs=0.2, phs=wgeo.phase(), form=DFT_FORMS[form])
    if corr:
        kw.update(correlationCentroiding=1, corrPattern=W.reference_images(wgeo), parabolicFit=5)
    if noise:
        kw.update(sig=1000.0, readnoise=0.5, addPoisson=1, noiseFloor=0.1, threshType=0, relativeFloor=True, seed=9)
    c = Case(nsubx, n, 64, 16, **kw)
    oc, ob, _ = c.run_both()
    act = _active(c)
    b = c.bimg.reshape(-1, 16, 16)[act]
    if noise:
        diff = np.abs(b - ob[act])
        diffpix = diff > 1e-3 * np.maximum(1, np.abs(ob[act]))
        say("noisy pixels differing: %d of %d" % (diffpix.sum(), diffpix.size))
        if "x1" in form:
            # one rounding to 16 bits moves the photon rate by up to _img_tol of the maximum, so
            # more pixels draw a neighbouring count from the same normal: bound their share, and
            # each such pixel differs by one or two photons (two where the rate is hundreds of
            # photons), plus the shift of the relative floor when the brightest pixel moves
            # (most of them by the floor shift alone: when the brightest pixel of a sub-aperture
            # moves by a photon, every pixel above the floor moves by a tenth of one)
            assert diffpix.sum() <= 2e-2 * diffpix.size
            if diffpix.any():
                say("largest difference of those pixels: %.3g" % diff[diffpix].max())
                assert diff[diffpix].max() <= 2.5
        else:
            assert diffpix.sum() <= max(3, 1e-3 * diffpix.size)
        clean = ~diffpix.reshape(diffpix.shape[0], -1).any(axis=1)
        d = np.abs(c.cents.reshape(-1, 2)[act] - oc[act])[clean]
        assert np.sqrt((d ** 2).mean()) < 1e-3
        c.free()
        return
    d = report("dft32-%s-%s" % (name, form), c.cents, oc, act)
    assert np.sqrt((d ** 2).mean()) < 1e-3
    assert d.max() < (1e-4 if _strict(form) else _max_bound(form))
    np.testing.assert_allclose(b, ob[act], rtol=0, atol=_img_tol(form) * ob[act].max())
    assert np.all(c.cents.reshape(-1, 2)[~act] == 0)
    c.free()


@pytest.mark.parametrize("m", [45, 83])
def test_interpolated_phase_grid(m):
    """Phase on an m x m grid (not nsubx*n = 64): interpolatePhase (bicubic) onto the 64 x 64
    grid before imaging, against the oracle's FP64 restatement."""
    say("case interpolatePhase m=%d" % m)
    from dasp_amd import synth
    phs = synth.phase_screen(m, 4.2, 0.15, 30.0, 77, wavelength_scale=500. / 640.) * 0.7
    c = Case(8, 8, 16, 8, obs=0.25, phs=phs.astype(np.float32))
    oc, ob, _ = c.run_both()
    act = _active(c)
    d = report("interp m=%d" % m, c.cents, oc, act)
    assert d.max() < 1e-4
    b = c.bimg.reshape(-1, 8, 8)[act]
    np.testing.assert_allclose(b, ob[act], rtol=0, atol=2e-5 * ob[act].max())
    c.free()


@pytest.mark.parametrize("seed,stream,n", [(1, 0, 1), (0, 1, 33), (7, 3, 200001), (11, 2 ** 40 + 5, 409600)])
def test_device_noise_stream_matches_host(seed, stream, n):
    """The device noise generator (dasp_cent_run's readout noise, WFSBatch.draw_noise) against the
    host generator: the same PCG32 integers, so equal values except where the device and host
    double-precision log/sqrt/cos/sin round to different floats (at most a few ulp, rarely)."""
    import torch
    from dasp_amd.cmod import cent
    say("case device normals seed=%d stream=%d n=%d" % (seed, stream, n))
    host = cent.drawNormals(seed, stream, n)
    dev = torch.full((n + 1,), 123.0, dtype=torch.float32, device="cuda:0")
    cent.drawNormalsDevice(seed, stream, n, dev.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    got = dev.cpu().numpy()
    assert got[n] == 123.0   # nothing written past n
    got = got[:n]
    diff = got != host
    say("values differing from the host: %d of %d" % (diff.sum(), n))
    assert diff.sum() <= 2 + n // 100000
    np.testing.assert_allclose(got, host, rtol=4 * np.finfo(np.float32).eps, atol=1e-30)


@pytest.mark.parametrize("form", ["fused", "fused_noise", "dft16_fp16x1", "dft32_fp16x1_corr"])
def test_launch_grid_tuner_invariance(form):
    """dasp_cent_set_grid / dasp_cent_tune_grid: every item is computed the same way whatever the
    launch grid, so the slopes of a multi-frame device run are bit for bit those of the default
    grid under capped, uncapped and tuned grids, for the image kernel and the correlation stage."""
    import torch
    from dasp_amd.cmod import cent
    say("case launch grid tuner %s" % form)
    kw = {}
    if form == "dft32_fp16x1_corr":
        wgeo = W.WFS("dft32", 12, 32, 64, 16, 8.0, 0.2, 0.15, 25.0)
        kw = dict(correlationCentroiding=1, corrPattern=W.reference_images(wgeo), parabolicFit=5)
        c = Case(12, 32, 64, 16, obs=0.2, phs=wgeo.phase(), form=4, **kw)
    else:
        wgeo = W.WFS("dft16", 11, 16, 32, 16, 8.0, 0.2, 0.15, 25.0)
        if form == "fused_noise":
            kw = dict(sig=1000.0, readnoise=0.5, addPoisson=1, noiseFloor=0.1, threshType=0, relativeFloor=True)
        c = Case(11, 16, 32, 16, obs=0.2, phs=wgeo.phase(), form=4 if form == "dft16_fp16x1" else 10, **kw)
    h = c.cc.centstruct
    nframes = 5
    frames = np.stack([wgeo.phase(f) for f in range(nframes)])
    dev = torch.device("cuda:0")
    ph = torch.from_numpy(frames).to(dev)
    S = c.nsubx * c.nsubx
    nact, npup, csz = cent.info(h)
    npix = S * c.nimg * c.nimg * nframes
    zp = torch.from_numpy(np.random.default_rng(7).standard_normal(npix).astype(np.float32)).to(dev)
    zr = torch.from_numpy(np.random.default_rng(8).standard_normal(npix).astype(np.float32)).to(dev)
    st = torch.cuda.current_stream().cuda_stream

    def run():
        ce = torch.full((nframes, S, 2), -9.0, dtype=torch.float32, device=dev)
        cent.runDevice(h, ph.data_ptr(), npup * npup, nframes, ce.data_ptr(), 0, 0, zp.data_ptr(), zr.data_ptr(), st)
        torch.cuda.synchronize()
        return ce.cpu().numpy()

    cent.setGrid(h, 0, 0)
    want = run()
    for gi, gc in [(8, 8), (256, 16), (-1, -1), (4096, 2048)]:
        cent.setGrid(h, gi, gc)
        assert cent.getGrid(h) == (gi, gc)
        np.testing.assert_array_equal(run(), want)
    ce = torch.zeros((nframes, S, 2), dtype=torch.float32, device=dev)
    gi, gc, ms = cent.tuneGrid(h, ph.data_ptr(), npup * npup, nframes, ce.data_ptr(), zp.data_ptr(), zr.data_ptr(),
                               reps=2, stream=st)
    say("tuned grids %d / %d, %.4f ms" % (gi, gc, ms))
    assert ms > 0 and cent.getGrid(h) == (gi, gc)
    np.testing.assert_array_equal(ce.cpu().numpy(), want)
    np.testing.assert_array_equal(run(), want)
    c.free()


@pytest.mark.parametrize("form", ["fused", "dft16_fp16", "dft16_bf16_fasttrig", "dft32_fp16_corr"])
def test_run_device_zeroes_inactive_slopes(form):
    """The image kernels zero the inactive sub-apertures' slopes of every frame themselves (no
    separate launch): an output buffer filled with garbage reads 0 there after a device run, and
    the active slopes equal the legacy run's."""
    import torch
    from dasp_amd.cmod import cent
    say("case inactive slopes %s" % form)
    if form == "dft32_fp16_corr":
        wgeo = W.WFS("dft32", 12, 32, 64, 16, 8.0, 0.2, 0.15, 25.0)
        c = Case(12, 32, 64, 16, obs=0.2, phs=wgeo.phase(), form=2, correlationCentroiding=1,
                 corrPattern=W.reference_images(wgeo), parabolicFit=5)
    else:
        wgeo = W.WFS("dft16", 11, 16, 32, 16, 8.0, 0.2, 0.15, 25.0)
        code = {"fused": 10, "dft16_fp16": 2, "dft16_bf16_fasttrig": 11}[form]
        c = Case(11, 16, 32, 16, obs=0.2, phs=wgeo.phase(), form=code)
    nframes = 3
    frames = np.stack([wgeo.phase(f) for f in range(nframes)])
    want = []
    for f in range(nframes):
        c.phs[:] = frames[f]
        c.cc.run(0)
        want.append(c.cents.copy())
    dev = torch.device("cuda:0")
    ph = torch.from_numpy(frames).to(dev)
    S = c.nsubx * c.nsubx
    ce = torch.full((nframes, S, 2), -9.0, dtype=torch.float32, device=dev)
    bi = torch.zeros((nframes, S, c.nimg, c.nimg), dtype=torch.float32, device=dev)
    nact, npup, csz = cent.info(c.cc.centstruct)
    ci = torch.zeros((nframes, S, csz, csz), dtype=torch.float32, device=dev)
    cent.runDevice(c.cc.centstruct, ph.data_ptr(), npup * npup, nframes, ce.data_ptr(), bi.data_ptr(),
                   corrimg_ptr=ci.data_ptr(), stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    got = ce.cpu().numpy().reshape(nframes, -1, 2)
    act = _active(c)
    assert np.all(got[:, ~act] == 0)
    np.testing.assert_array_equal(got[:, act], np.stack(want).reshape(nframes, -1, 2)[:, act])
    c.free()


@pytest.mark.parametrize("form", [10, 12], ids=["fused_fasttrig", "dft16_fp16_fasttrig"])
def test_degenerate_geometries(form):
    """Edge cases: every sub-aperture inactive (0.99 obstruction: slopes all 0, no image kernel),
    a single sub-aperture (nsubx 1: seven of the eight XCD slices empty), and a device step of
    zero frames (rejected with an error, buffers untouched)."""
    import torch
    from dasp_amd.cmod import cent
    say("case degenerate geometries form %d" % form)
    c = Case(2, 16, 32, 16, obs=0.99, form=form)
    assert c.cc.subflag.sum() == 0
    c.cents[:] = -9.0
    oc, _, _ = c.run_both()
    assert np.all(c.cents == 0) and np.all(oc == 0)
    c.free()
    w = W.WFS("one", 1, 16, 32, 16, 0.5, 0.0, 0.15, 25.0)
    c = Case(1, 16, 32, 16, obs=0.0, phs=w.phase(), form=form)
    oc, ob, _ = c.run_both()
    act = _active(c)
    assert act.sum() == 1
    assert report("one subaperture", c.cents, oc, act).max() < 1e-4
    dev = torch.device("cuda:0")
    ph = torch.from_numpy(c.phs).to(dev)
    ce = torch.full((1, 1, 2), -9.0, dtype=torch.float32, device=dev)
    with pytest.raises(cent.error):
        cent.runDevice(c.cc.centstruct, ph.data_ptr(), 256, 0, ce.data_ptr(),
                       stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert np.all(ce.cpu().numpy() == -9.0)
    c.free()
