"""GPU: the reconstructor step of science/tomoRecon.py on the device (dasp_recon_*), fed by engine
slopes that stay in HBM, against a float64 NumPy restatement: data = slopes gathered in the
reconstructor's order (x slopes of every WFS, then y slopes), acts = decay*acts - gains*(data.R)."""
import numpy as np
import pytest

from dasp_amd import workloads as W
from dasp_amd.science.tomorecon import DeviceRecon, cent_index
from tests._cases import Case, say

pytestmark = pytest.mark.gpu


def _numpy_step(slopes, index, R, nacts, layout, gains, decay, acts):
    data = slopes[:, index].astype(np.float64)
    tmp = data @ (R.astype(np.float64) if layout == 0 else R.astype(np.float64).T)
    return decay * acts - gains * tmp[:, :nacts]


@pytest.mark.parametrize("layout,frames", [(0, 1), (0, 3), (0, 11), (1, 1), (1, 5)])
def test_recon_step_matches_numpy(layout, frames):
    import torch
    say("case recon layout %d frames %d" % (layout, frames))
    rng = np.random.default_rng(layout * 10 + frames)
    flags = [(rng.random((n, n)) > 0.3).astype(np.int32) for n in (20, 8, 13)]
    ncents = 2 * sum(int(f.sum()) for f in flags)
    nmodes, nacts = 1031, 1025                  # nacts not a multiple of 4: the scalar column tail
    shape = (ncents, nmodes) if layout == 0 else (nmodes, ncents)
    R = (rng.standard_normal(shape) / np.sqrt(ncents)).astype(np.float32)
    gains = (0.2 + 0.3 * rng.random(nacts)).astype(np.float32)
    rec = DeviceRecon(flags, torch.from_numpy(R).to("cuda:0"), nacts=nacts, gains=gains, decay=0.99,
                      layout=layout, frames=frames)
    assert rec.ncents == ncents
    slopes = rng.standard_normal((frames, rec.slope_len)).astype(np.float32)
    acts0 = rng.standard_normal((frames, nacts)).astype(np.float32)
    rec.slopes.copy_(torch.from_numpy(slopes))
    rec.acts.copy_(torch.from_numpy(acts0))
    rec.run()
    torch.cuda.synchronize()
    got = rec.acts.cpu().numpy()
    want = _numpy_step(slopes, rec.index, R, nacts, layout, gains.astype(np.float64), 0.99, acts0)
    err = np.abs(got - want).max() / np.abs(want).max()
    say("  max relative error %.3g" % err)
    assert err < 2e-6
    rec.run()                                   # the same again: deterministic summation order
    torch.cuda.synchronize()
    want2 = _numpy_step(slopes, rec.index, R, nacts, layout, gains.astype(np.float64), 0.99, got.astype(np.float64))
    assert np.abs(rec.acts.cpu().numpy() - want2).max() / np.abs(want2).max() < 2e-6
    rec.free()


def test_engine_slopes_feed_recon_in_hbm():
    """Two WFS write their slopes straight into the reconstructor's buffer (device runs); the
    step's actuators equal NumPy's on the same slopes copied back."""
    import torch
    from dasp_amd.cmod import cent
    say("case engine -> recon in HBM")
    ws = [W.WFS("r1", 12, 16, 32, 16, 8.0, 0.2, 0.15, 25.0), W.WFS("r2", 8, 8, 16, 8, 4.2, 0.25, 0.2, 30.0)]
    cases = [Case(w.nsubx, w.n, w.nfft, w.nimg, obs=w.obs) for w in ws]
    flags = [c.cc.subflag for c in cases]
    ncents = 2 * sum(int(f.sum()) for f in flags)
    rng = np.random.default_rng(3)
    nacts = 300
    R = (rng.standard_normal((ncents, nacts)) * 0.1).astype(np.float32)
    rec = DeviceRecon(flags, torch.from_numpy(R).to("cuda:0"), gains=0.5, decay=1.0)
    st = torch.cuda.current_stream().cuda_stream
    for k, (w, c) in enumerate(zip(ws, cases)):
        ph = torch.from_numpy(w.phase()).to("cuda:0")
        out = rec.slope_buffer(k)
        cent.runDevice(c.cc.centstruct, ph.data_ptr(), w.npup * w.npup, 1, out.data_ptr(), stream=st)
    rec.run()
    torch.cuda.synchronize()
    slopes = rec.slopes.cpu().numpy()
    assert np.abs(slopes).max() > 0
    want = _numpy_step(slopes, cent_index(flags, rec.offsets), R, nacts, 0, 0.5, 1.0, np.zeros((1, nacts)))
    err = np.abs(rec.acts.cpu().numpy() - want).max() / np.abs(want).max()
    say("  max relative error %.3g" % err)
    assert err < 2e-6
    rec.free()
    for c in cases:
        c.free()
