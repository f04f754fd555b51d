"""Host logic of the generateNext mirror (science/wfscent.py:630-733) with a stand-in centroid
object: readout cadence of the integration cycle, tip-tilt removal over all nsubx^2 entries,
covariance sums, zeroOutput. No GPU."""
import numpy as np

from dasp_amd.science.wfscent import wfscent


class FakeCentroid:
    """Records runCalc calls; each call writes a distinct slope pattern."""

    def __init__(self, nsubx=4, nint=1, nlat=0, subtractTipTilt=0):
        self.nsubx = nsubx
        self.nsubaps = nsubx * nsubx
        self.nintegrations, self.nlatency = nint, nlat
        self.curInteg = 0
        self.imageOnly = 0
        self.device = None
        self.subtractTipTilt = subtractTipTilt
        self.phase = np.zeros((8, 8), np.float32)
        self.outputData = np.zeros((nsubx, nsubx, 2), np.float32)
        self.calls = 0

    def runCalc(self, control):
        self.calls += 1
        rng = np.random.default_rng(self.calls)
        self.outputData[:] = rng.normal(size=self.outputData.shape) + 0.3


def test_readout_cadence_and_phase_copy():
    w = FakeCentroid(nint=2, nlat=1)
    m = wfscent(w)
    valid = []
    for k in range(7):
        phs = np.full((8, 8), k, np.float32)
        m.generateNext(phs)
        valid.append(m.dataValid)
        assert np.all(w.phase == k)          # copied when the parent's array is another one
    assert valid == [0, 0, 1, 0, 0, 1, 0]
    assert w.calls == 7                      # runCalc on every step; the engine reads out on the third


def test_tip_tilt_removal_uses_all_entries():
    w = FakeCentroid(subtractTipTilt=1)
    m = wfscent(w)
    m.generateNext(w.phase)
    rng = np.random.default_rng(1)
    raw = rng.normal(size=w.outputData.shape) + 0.3
    want = raw.copy()
    want[:, :, 0] -= raw[:, :, 0].sum() / 16
    want[:, :, 1] -= raw[:, :, 1].sum() / 16
    np.testing.assert_allclose(w.outputData, want, rtol=1e-6, atol=1e-6)
    # calibration source: no removal unless subtractTipTilt == -1
    w2 = FakeCentroid(subtractTipTilt=1)
    m2 = wfscent(w2, {"cal_source": 1})
    m2.generateNext(w2.phase)
    np.testing.assert_allclose(w2.outputData, raw.astype(np.float32), rtol=1e-6)


def test_covariance_sums_and_zero_output():
    w = FakeCentroid()
    m = wfscent(w, {"calcCovariance": 1})
    outs = []
    for _ in range(3):
        m.generateNext(w.phase)
        outs.append(w.outputData.copy())
    np.testing.assert_allclose(w.outSum, sum(outs), rtol=1e-6)
    np.testing.assert_allclose(w.outSquare, sum(o ** 2 for o in outs), rtol=1e-6)
    assert w.outN == 3
    z = wfscent(FakeCentroid(), {"zeroOutput": 1})
    z.wfscentObj.outputData[:] = 5
    z.generateNext(z.wfscentObj.phase)
    assert z.dataValid == 1 and np.all(z.wfscentObj.outputData == 0) and z.wfscentObj.calls == 0
