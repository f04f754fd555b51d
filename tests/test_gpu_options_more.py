"""GPU parity for the remaining reference options: pre-binning before the spot convolution
(prebinImage, cmod/src/centmodule.c:1583-1626), difference correlation with flux normalisation and
the fallback loop (correlationCentroiding 3 and 4, :1203-1338), phase+amplitude input (:488-497,
:1834-1853) and image input cut out by subapLocation (:2009-2030). All run through the drop-in
centcmod contract and are compared with the FP64 oracle on the same inputs."""
import numpy as np
import pytest

from dasp_amd import workloads as W
from oracle import centoracle as O
from tests._cases import Case, report, say

pytestmark = pytest.mark.gpu


def _active(c):
    return c.cc.subflag.ravel() == 1


def _compare(c, tag, oc, ob, ocorr=None):
    act = _active(c)
    nimg = c.nimg
    b = c.bimg.reshape(-1, nimg, nimg)[act]
    np.testing.assert_allclose(b, ob[act], rtol=0, atol=2e-5 * max(1e-30, np.abs(ob[act]).max()))
    assert report(tag, c.cents, oc, act).max() < 1e-4
    if ocorr is not None and c.corrimg is not None:
        K = c.corrimg.shape[-1]
        np.testing.assert_allclose(c.corrimg.reshape(-1, K, K)[act], ocorr[act], rtol=0,
                                   atol=1e-4 * max(1e-30, np.abs(ocorr[act]).max()))


PREBIN = {
    # name: (nfft, clip, nimg, spot size, preBinningFactor): psfsize between clip and paddedsize
    "pbf2_p16": (32, 16, 8, 16, 2),      # 32 -> 16 binned, convolved at 16
    "pbf4_p16": (32, 16, 8, 16, 4),      # 32 -> 8 binned, zero-padded to 16 before convolution
    "pbf2_p32_clip16": (64, 16, 8, 32, 2),
}


@pytest.mark.parametrize("name", sorted(PREBIN))
def test_prebinning_with_spot(name):
    nfft, clip, nimg, P, pbf = PREBIN[name]
    nsubx, n = 6, 16
    say("case prebin %s" % name)
    wgeo = W.WFS("pb", nsubx, n, nfft, nimg, 8.0, 0.2, 0.15, 25.0, clipsize=clip, spot=True, spot_size=P,
                 sig=1000.0)
    c = Case(nsubx, n, nfft, nimg, clip=clip, obs=0.2, sig=1000.0, spotpsf=wgeo.spots(), preBinningFactor=pbf,
             phs=wgeo.phase())
    assert c.params.sizes() == (P, nfft, pbf)
    oc, ob, _ = c.run_both()
    _compare(c, "prebin " + name, oc, ob)
    c.free()


@pytest.mark.parametrize("mode", [3, 4])
def test_difference_correlation_modes(mode):
    """correlationCentroiding 3 takes diffCorrelation's fallback loop (sum |d|, squared), 4 the
    flux-normalised sum of d^2; both then find the minimum and fit (parabolicFit 3)."""
    nsubx, n, nfft, nimg = 6, 8, 16, 8
    say("case difference correlation mode %d" % mode)
    wgeo = W.WFS("dc", nsubx, n, nfft, nimg, 4.2, 0.25, 0.2, 30.0)
    ref = W.reference_images(wgeo).astype(np.float32)
    for fit in (0, 3):
        c = Case(nsubx, n, nfft, nimg, obs=0.25, phs=wgeo.phase() * 0.5, correlationCentroiding=mode,
                 corrPattern=ref, parabolicFit=fit)
        oc, ob, ocorr = c.run_both()
        _compare(c, "diffcorr mode %d fit %d" % (mode, fit), oc, ob, ocorr)
        c.free()


def test_phase_and_amplitude_input():
    """A (2, npup, npup) phase+amplitude array: the field is amp * pup * exp(i(phi - t)) and the
    normalisation uses the amplitude-weighted pupil total."""
    nsubx, n, nfft, nimg = 8, 8, 16, 8
    say("case phase+amplitude")
    w = W.WFS("pa", nsubx, n, nfft, nimg, 4.2, 0.25, 0.2, 30.0)
    npup = nsubx * n
    rng = np.random.default_rng(4)
    yy, xx = np.mgrid[0:npup, 0:npup] / npup
    amp = (0.6 + 0.4 * np.cos(2 * xx + yy) + 0.05 * rng.random((npup, npup))).astype(np.float32)
    phs = np.ascontiguousarray(np.stack([w.phase() * 0.5, amp]).astype(np.float32))
    c = Case(nsubx, n, nfft, nimg, obs=0.25, phs=phs, sig=100.0)
    oc, ob, _ = c.run_both()
    _compare(c, "phase+amp", oc, ob)
    c.free()


def test_image_input_subap_location():
    """runSlope on a 2-D camera image with a subapLocation table: each window (some smaller than
    nimg) is copied to the top-left of the sub-aperture image, the rest keeps the previous image."""
    nsubx, nimg = 6, 8
    say("case runSlope with subapLocation")
    rng = np.random.default_rng(5)
    big = (rng.random((nsubx * nimg, nsubx * nimg)) * 50).astype(np.float32)   # (nsubx*nimg)^2, :3118-3157
    S = nsubx * nsubx
    loc = np.zeros((S, 4), np.int32)
    for s in range(S):
        r0, c0 = (s // nsubx) * nimg + (s % 3), (s % nsubx) * nimg + (s % 2)
        h, w = nimg - (s % 3), nimg - 2 * (s % 2)
        loc[s] = [r0, r0 + h, c0, c0 + w]
    c = Case(nsubx, 8, 16, nimg, obs=0.25, noiseFloor=2.0, inputImage=big, subapLocation=loc)
    c.bimg[:] = 3.0                                  # pixels outside a window keep this value
    before = c.bimg.copy()
    c.cc.runSlope(0)
    oc, ob, _ = O.run_image(c.params, big, bimg_in=before, subap_location=loc)
    _compare(c, "runSlope subapLocation", oc, ob)
    c.free()
