"""GPU: every update code of cent.update (cmod/src/centmodule.c:2573-2856, codes :109-130) applied
to a live handle, then a run compared with the FP64 oracle given the same state.

Each case builds a small WFS (6x6, n=8, nfft=16, nimg=8), runs it once, applies one update (or a
short sequence) through centcmod.update exactly as the reference's callers do (util/centroid.py
:397-402, :623-627, :1180-1202, :1334-1357, :1523-1526, :1546-1679), mirrors the change in the
oracle's parameters, runs again and compares. Bars as in the parity tests: noiseless slopes within
1e-4 px and images within 1e-4 of their peak; noisy runs bound the pixels that round differently.
"""
import numpy as np
import pytest

from dasp_amd import workloads as W
from dasp_amd.util import centcmod as C
from oracle import centoracle as O
from tests._cases import Case, report, say

pytestmark = pytest.mark.gpu

NSUBX, N, NFFT, NIMG = 6, 8, 16, 8
S = NSUBX * NSUBX
GEO = W.WFS("upd", NSUBX, N, NFFT, NIMG, 4.2, 0.25, 0.2, 30.0)


def _case(**kw):
    return Case(NSUBX, N, NFFT, kw.pop("nimg", NIMG), obs=0.25, phs=GEO.phase() * 0.5, **kw)


def _check(c, tag, noisy=False):
    oc, ob, ocorr = c.run_both()
    act = c.cc.subflag.ravel() == 1
    nimg = c.nimg
    b = c.bimg.reshape(-1, nimg, nimg)[act]
    if noisy:
        diffpix = np.abs(b - ob[act]) > 1e-3 * np.maximum(1, np.abs(ob[act]))
        say("%s: noisy pixels differing %d of %d" % (tag, diffpix.sum(), diffpix.size))
        assert diffpix.sum() <= 3
        clean = ~diffpix.reshape(diffpix.shape[0], -1).any(axis=1)
        d = np.abs(c.cents.reshape(-1, 2)[act] - oc[act])
        assert d[clean].max() < 1e-4
        return oc, ob
    np.testing.assert_allclose(b, ob[act], rtol=0, atol=1e-4 * max(1e-30, np.abs(ob[act]).max()))
    if not c.params.imageOnly:
        assert report(tag, c.cents, oc, act).max() < 1e-4
    if ocorr is not None and c.corrimg is not None:
        K = c.corrimg.shape[-1]
        np.testing.assert_allclose(c.corrimg.reshape(-1, K, K)[act], ocorr[act], rtol=0,
                                   atol=1e-4 * max(1e-30, np.abs(ocorr[act]).max()))
    return oc, ob


def _update(c, code, value, **oracle):
    c.cc.update(code, value)
    for k, v in oracle.items():
        assert hasattr(c.params, k), k
        setattr(c.params, k, v)


def _spot(P, seed=1):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:P, 0:P] - (P - 1) / 2.0
    sp = np.exp(-(x * x / 3.0 + y * y / 1.2)) + 0.01 * rng.random((P, P))
    return (sp / sp.sum()).astype(np.float32)


def _caldata(steps=21):
    st = np.linspace(-1.5, 1.5, steps).astype(np.float32)
    cd = np.empty((2, NSUBX, NSUBX, steps), np.float32)
    for d in range(2):
        for s in range(S):
            cd[d, s // NSUBX, s % NSUBX] = st * (1.1 + 0.01 * s) + 0.02 * d   # increasing tables
    cb = np.empty((2, NSUBX, NSUBX, 2), np.int32)
    cb[..., 0], cb[..., 1] = 2, steps - 3
    return cd, cb, st


def case_calsource():
    c = _case(sig=500.0, readnoise=0.5, addPoisson=1, useBrightest=10, noiseFloor=0.5)
    _check(c, "calsource 0", noisy=True)
    _update(c, C.CALSOURCE, 1, calsource=1)        # no Poisson, no read noise, no pixel selection
    return c, False


def case_sig_scalar():
    c = _case(sig=1.0)
    _check(c, "sig 1")
    _update(c, C.SIG, 3.5, sig=3.5)
    return c, False


def case_sig_array_then_scalar():
    """A scalar SIG after an array SIG only sets sig: normalisation keeps the array (py_update
    :2596-2606 leaves sigArr in place; centroidsFromPhase reads sigArr[i] when it is set)."""
    arr = (np.random.default_rng(2).random(S) * 2 + 0.5).astype(np.float32)
    c = _case(sig=1.0)
    _update(c, C.SIG, arr, sigArr=arr)
    _check(c, "sig array")
    _update(c, C.SIG, 7.0, sig=7.0)                # oracle keeps sigArr, as the reference does
    return c, False


def case_addpoisson():
    c = _case(sig=800.0)
    _check(c, "no poisson")
    _update(c, C.ADDPOISSON, 1, addPoisson=1)
    return c, True


def case_ncen():
    c = _case()
    _update(c, C.NCEN, 6, ncen=6)
    return c, False


def case_readnoise_readbg_floor():
    c = _case(sig=800.0)
    _update(c, C.READNOISE, 0.7, readnoise=0.7)
    _update(c, C.READBG, 0.3, readbg=0.3)
    _update(c, C.NOISEFLOOR, 1.0, noiseFloor=1.0)
    return c, True


def case_skybrightness():
    c = _case(sig=10.0)
    _update(c, C.SKYBRIGHTNESS, 0.05, skybrightness=0.05)
    _check(c, "sky scalar")
    sky = (np.random.default_rng(4).random((S, NIMG, NIMG)) * 0.02).astype(np.float32)
    _update(c, C.SKYBRIGHTNESS, sky, skybrightness=0.0, skybrightnessArr=sky)
    _check(c, "sky array")
    _update(c, C.SKYBRIGHTNESS, 0.01, skybrightness=0.01, skybrightnessArr=None)
    return c, False


def case_pxlpower():
    c = _case()
    _update(c, C.PXLPOWER, 2.0, pxlPower=2.0)
    _check(c, "power 2")
    _update(c, C.PXLPOWER, 1.5, pxlPower=1.5)
    return c, False


def case_spotpsf():
    c = _case(sig=100.0)
    sp = _spot(NFFT)
    _update(c, C.SPOTPSF, sp, spotpsf=sp)
    _check(c, "spot 2-D")
    sp4 = np.stack([_spot(NFFT, k) for k in range(S)]).reshape(NSUBX, NSUBX, NFFT, NFFT)
    _update(c, C.SPOTPSF, sp4, spotpsf=sp4)
    _check(c, "spot 4-D")
    _update(c, C.SPOTPSF, None, spotpsf=None)
    return c, False


def case_nthreads():
    c = _case()
    _update(c, C.NTHREADS, 5)                      # the host pool of the reference; no effect here
    return c, False


def case_opticalbinning():
    c = _case()
    _update(c, C.OPTICALBINNING, 1, opticalBinning=1)
    return c, False


def case_centweight():
    w2 = (np.random.default_rng(6).random((NIMG, NIMG)) + 0.5).astype(np.float32)
    c = _case()
    _update(c, C.CENTWEIGHT, w2, centWeight=w2)
    _check(c, "centweight 2-D")
    _update(c, C.CENTWEIGHT, None, centWeight=None)
    return c, False


def _corr_case(**kw):
    ref = W.reference_images(GEO).astype(np.float32)
    return _case(correlationCentroiding=1, corrPattern=ref, parabolicFit=3, **kw), ref


def case_correlation_switch_and_thresh():
    c, ref = _corr_case()
    _check(c, "corr 1")
    _update(c, C.CORRELATIONCENTROIDING, 0, correlationCentroiding=0)   # fit on the image itself
    _check(c, "corr off")
    _update(c, C.CORRELATIONCENTROIDING, 1, correlationCentroiding=1)
    _update(c, C.CORRTHRESH, 0.3, corrThresh=0.3)
    return c, False


def case_corrpattern():
    c, ref = _corr_case()
    ref2 = np.roll(ref, 1, axis=-1).copy()
    hc2 = O.transform_psf(ref2)
    _update(c, C.CORRPATTERN, hc2, corrPattern=hc2)
    return c, False


def case_caldata():
    c = _case()
    cd, cb, st = _caldata()
    _update(c, C.CALDATA, (cd, cb, st), calData=cd, calBounds=cb, calSteps=st)
    _check(c, "caldata")
    _update(c, C.CALDATA, None, calData=None, calBounds=None, calSteps=None)
    return c, False


def case_calcoeff():
    c = _case()
    co = np.zeros((NSUBX, NSUBX, 2, 3), np.float32)
    co[..., 0] = 0.01
    co[..., 1] = 1.2
    co[..., 2] = -0.05
    _update(c, C.CALCOEFF, co, calCoeff=co)
    _check(c, "calcoeff")
    _update(c, C.CALCOEFF, None, calCoeff=None)
    return c, False


def case_refcents():
    c = _case()
    rc = (np.random.default_rng(8).random((NSUBX, NSUBX, 2)) * 0.2 - 0.1).astype(np.float32)
    _update(c, C.REFCENTS, rc, refCents=rc)
    _check(c, "refcents")
    _update(c, C.REFCENTS, None, refCents=None)
    return c, False


def case_usebrightest():
    c = _case(sig=50.0)
    _update(c, C.USEBRIGHTEST, 12, useBrightest=12)
    _check(c, "brightest 12")
    ub = np.random.default_rng(9).integers(-20, 20, S).astype(np.int32)
    _update(c, C.USEBRIGHTEST, ub, useBrightest=0, useBrightestArr=ub)
    return c, False


CASES = {k[5:]: v for k, v in dict(globals()).items() if k.startswith("case_")}


@pytest.mark.parametrize("name", sorted(CASES))
def test_update_code(name):
    say("case update %s" % name)
    c, noisy = CASES[name]()
    _check(c, name, noisy)
    c.free()


def test_update_integsteps():
    """INTEGSTEPS (:2822-2829) shortens the integration: with nintegrations 2 -> 1 every call reads out."""
    say("case update integsteps")
    c = _case(nintegrations=2)
    outs = []
    for k in range(4):
        if k == 2:
            c.cc.update(C.INTEGSTEPS, 1)
        c.cents[:] = -7
        c.cc.run(0)
        outs.append(bool(np.all(c.cents != -7)))
    assert outs == [False, True, True, True]
    c.free()


def test_update_outputphasearr():
    """OUTPUTPHASEARR (:2832-2848): with a phase grid to interpolate, the interpolated phase of every
    active sub-aperture is written into the given array (the reference's verification output)."""
    say("case update outputphasearr")
    m = 45
    rng = np.random.default_rng(10)
    yy, xx = np.mgrid[0:m, 0:m] / (m - 1.0)
    phs = (np.sin(3 * xx) * np.cos(2 * yy) + 0.1 * rng.standard_normal((m, m))).astype(np.float32)
    c = Case(NSUBX, N, NFFT, NIMG, obs=0.25, phs=phs)
    out = np.full((NSUBX * N, NSUBX * N), -99.0, np.float32)
    c.cc.update(C.OUTPUTPHASEARR, out)
    _check(c, "interpolated run")
    want = O.interpolate_phase(phs, NSUBX, N)
    act = (c.cc.subflag == 1)
    mask = np.kron(act, np.ones((N, N), bool))
    np.testing.assert_allclose(out[mask], want[mask], rtol=0, atol=2e-5)
    assert np.all(out[~mask] == -99.0)
    c.cc.update(C.OUTPUTPHASEARR, None)
    c.free()
