"""GPU: the set-up routines of the centroid mirror (util/centroid.py) over the engine.

- runSlopeCalc (:547-558): image input through the engine's runSlope, against the oracle, and a
  reference taken from the camera image (takeReference with cameraInput) zeroing the slopes;
- takeCentWeight (:1661-1679): the flat-wavefront image as CoG weighting, against the oracle's
  weighted CoG with that weight;
- calibrateSHS (:1492-1504): the tilt sweep and its inversion. Identical-pupil calibration runs in
  Python after the engine; per-sub-aperture tables (with a spot kernel) and polynomials go to the
  engine through CALDATA / CALCOEFF. In every case a tilt of one of the sweep's steps must come back
  as that step (the calibration inverts the sweep), to the accuracy the method allows.
"""
import numpy as np
import pytest

from dasp_amd import workloads as W
from dasp_amd.util.centroid import centroid
from dasp_amd.util.tel import obstructed_pupil
from oracle import centoracle as O
from tests._cases import say

pytestmark = pytest.mark.gpu

NSUBX, N, NFFT, NIMG = 6, 8, 16, 8


def _make(**kw):
    pup = obstructed_pupil(NSUBX, N, 0.25, 0.5)
    c = centroid(NSUBX, pup=pup, phasesize=N, fftsize=NFFT, nimg=NIMG, **kw)
    c.initMem()
    c.initialiseCmod()
    return c


def _params(c, **kw):
    cc = c.centcmod
    return O.Params(NSUBX, N, NFFT, c.clipsize, c.nimg, c.ncen, cc.pupfn.reshape(-1, N, N), cc.subflag,
                    cc.fracSubArea, sig=float(c.sig), **kw)


def test_run_slope_calc_image_input_and_camera_reference():
    say("case runSlopeCalc image input")
    src = _make()
    w = W.WFS("rs", NSUBX, N, NFFT, NIMG, 4.2, 0.25, 0.2, 30.0)
    src.phase[:] = w.phase() * 0.5
    src.runCalc({"cal_source": 1})
    img4 = src.cmodbimg.copy()
    camera = np.ascontiguousarray(img4.transpose(0, 2, 1, 3).reshape(NSUBX * NIMG, NSUBX * NIMG))
    src.closeCmod()
    c = _make(inputImage=camera.copy())   # floor 0: the cal_source 1 reference run skips the floor
    c.runSlopeCalc({"cal_source": 0})
    p = _params(c)
    act = c.centcmod.subflag.ravel() == 1
    oc, _, _ = O.run_image(p, camera)
    d = np.abs(c.outputData.reshape(-1, 2)[act] - oc[act]).max()
    say("  runSlopeCalc vs oracle: max|d| = %.3g px" % d)
    assert d < 1e-4
    ref = c.takeReference({"cal_source": 0}, cameraInput=camera)
    np.testing.assert_allclose(ref.reshape(-1, 2)[act], oc[act], rtol=0, atol=1e-4)
    c.runSlopeCalc({"cal_source": 0})
    assert np.abs(c.outputData).max() < 1e-6           # same image minus its own reference
    c.closeCmod()


def test_take_cent_weight():
    say("case takeCentWeight")
    c = _make(centWeight="make")
    c.takeCentWeight({"cal_source": 0})
    wgt = np.asarray(c.centWeight)
    assert wgt.shape == (NSUBX, NSUBX, NIMG, NIMG) and wgt.max() > 0
    w = W.WFS("cw", NSUBX, N, NFFT, NIMG, 4.2, 0.25, 0.2, 30.0)
    phs = (w.phase() * 0.5).astype(np.float32)
    c.phase[:] = phs
    c.runCalc({"cal_source": 0})
    oc, _, _ = O.run_phase(_params(c, centWeight=wgt), phs)
    act = c.centcmod.subflag.ravel() == 1
    d = np.abs(c.outputData.reshape(-1, 2)[act] - oc[act]).max()
    say("  weighted CoG vs oracle: max|d| = %.3g px" % d)
    assert d < 1e-4
    c.closeCmod()


def _tilt(c, step, axis):
    npup = NSUBX * N
    t = (-np.arange(npup) * step).astype(np.float32)
    return np.ascontiguousarray(np.broadcast_to(t[None, :] if axis == 0 else t[:, None], (npup, npup)))


@pytest.mark.parametrize("kind", ["identical", "unique_tables", "polynomial"])
def test_calibrate_shs_inverts_the_sweep(kind):
    say("case calibrateSHS %s" % kind)
    steps = 15
    kw = dict(linearSteps=steps, stepRangeFrac=0.5)
    if kind == "unique_tables":      # a spot kernel makes the calibration per sub-aperture, in the engine
        y, x = np.mgrid[0:NFFT, 0:NFFT] - (NFFT - 1) / 2.0
        spot = np.exp(-(x * x + y * y) / 2.0).astype(np.float32)
        kw.update(spotpsf=spot / spot.sum(), sig=100.0)
    if kind == "polynomial":
        kw.update(calNCoeff=4)
    c = _make(**kw)
    c.calibrateSHS({"cal_source": 1})
    act = c.centcmod.subflag == 1
    if kind == "identical":
        assert len(c.calDataDict) >= 2                  # fully and partly lit pupil types
    tol = {"identical": 1e-4, "unique_tables": 1e-4, "polynomial": 0.05}[kind]
    for k in range(steps // 3, 2 * steps // 3):
        st = float(c.calibrateSteps[k])
        for axis in (0, 1):
            c.phase[:] = _tilt(c, st, axis)
            c.runCalc({"cal_source": 0})
            got = np.asarray(c.outputData)[..., axis][act]
            err = np.abs(got - st).max()
            say("  step %d (%.4f) axis %d: max|calibrated - step| = %.3g" % (k, st, axis, err))
            assert err < tol
    c.closeCmod()
