"""GPU: the util/centroid + science/wfscent mirrors, host route (centcmod.run with numpy buffers,
as the reference) against the device route (phase, slopes, images and covariance sums in HBM):
same readout cadence, same noise streams, same slopes; reference slopes and the correlation
reference image taken the same way."""
import numpy as np
import pytest

from dasp_amd import workloads as W
from dasp_amd.science.wfscent import wfscent
from dasp_amd.util.centroid import new_centroid
from tests._cases import say

pytestmark = pytest.mark.gpu


def _make(device, **kw):
    w = W.SCAO8
    args = dict(nimg=8, sig=500., readnoise=0.5, addPoisson=1, noiseFloor=0.05, seed=3, device=device)
    args.update(kw)
    return new_centroid(w.nsubx, w.pupil(), w.n, w.nfft, **args)


def _host(x):
    return x.detach().cpu().numpy() if hasattr(x, "detach") else np.asarray(x)


def test_generate_next_host_and_device_routes_agree():
    import torch
    say("case wfscent host vs device, integration 2 + latency 1, tip-tilt, covariance")
    kw = dict(tstep=0.05, integtime=0.1, latency=0.05, subtractTipTilt=1)
    hc, dc = _make(None, **kw), _make("cuda:0", **kw)
    mh, md = wfscent(hc, {"calcCovariance": 1}), wfscent(dc, {"calcCovariance": 1})
    for k in range(6):
        phs = W.SCAO8.phase(k)
        mh.generateNext(phs.copy())
        md.generateNext(torch.from_numpy(phs).cuda())
        torch.cuda.synchronize()
        assert mh.dataValid == md.dataValid == (1 if k % 3 == 2 else 0)
        if mh.dataValid:
            d = np.abs(_host(dc.outputData) - hc.outputData).max()
            say("step %d: max|device - host| = %.3g px" % (k, d))
            assert d <= 1e-6
            assert abs(float(hc.outputData[:, :, 0].sum())) < 1e-3   # tip-tilt removed
    assert hc.outN == dc.outN == 2
    np.testing.assert_allclose(_host(dc.outSum), hc.outSum, rtol=0, atol=2e-6)
    np.testing.assert_allclose(_host(dc.outSquare), hc.outSquare, rtol=1e-5, atol=1e-6)
    hc.closeCmod()
    dc.closeCmod()


@pytest.mark.parametrize("device", [None, "cuda:0"])
def test_take_corr_image_and_reference(device):
    """takeCorrImage then takeReference with correlation + 3x3 parabolic fit: the flat wavefront
    measured against its own reference gives zero slopes, and a tilted one a non-zero slope of
    the same sign on both routes."""
    import torch
    say("case takeCorrImage + takeReference on %s" % (device or "host"))
    c = _make(device, correlationCentroiding=1, parabolicFit=3, readnoise=0., addPoisson=0, noiseFloor=0.)
    pat = c.takeCorrImage({"cal_source": 1})
    assert pat is not None and abs(pat.max() - 1.0) < 1e-6
    ref = c.takeReference({"cal_source": 0})
    assert ref.shape == (8, 8, 2) and np.abs(ref).max() < 1.0
    c.phase[:] = 0
    c.runCalc({"cal_source": 0})
    if device:
        torch.cuda.synchronize()
    act = c.centcmod.subflag == 1
    assert np.abs(_host(c.outputData)[act]).max() < 1e-5
    npup = c.nsubx * c.phasesize
    tilt = np.tile((np.arange(npup, dtype=np.float32) * 0.2)[None, :], (npup, 1))
    c.phase[:] = torch.from_numpy(tilt).cuda() if device else tilt
    c.runCalc({"cal_source": 0})
    if device:
        torch.cuda.synchronize()
    sx = _host(c.outputData)[act][:, 0]
    say("tilted: mean x slope %.4f px" % sx.mean())
    assert sx.mean() < -0.05                 # +x phase ramp -> -x slope (180-degree image rotation)
    c.closeCmod()
