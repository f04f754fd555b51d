"""GPU: the benchmark workloads at their full sizes, against the FP64 oracle.

- 80x80 Nyquist WFS with correlation (n32/nfft64/nimg16, 2560^2 phase, correlation mode 1 + 5x5
  parabolic fit): 4 frames through the device route with no image or correlation-image buffer
  (the engine's own scratch carries the images to the correlation stage), each of the FP32,
  FP16 and BF16 forms, every frame and every active sub-aperture against the oracle.
- 80x80 nfft32 (n16, 1280^2): the benchmark's call, 64 frames in one device run with no image
  buffer, every frame equal to a one-frame run of the same phase, and four frames (first, last,
  two between) against the oracle, for each form.
- The multi-WFS set (6 LGS 80x80 with per-subaperture spot kernels, 3 NGS 8x8 n32/nfft64/nimg8,
  1 truth 40x40) in one step graph, 2 frames: noiseless, every WFS against the oracle; with Poisson
  and read noise from the engine's noise streams, the pixels that round differently bounded and
  the other sub-apertures' slopes within 1e-4 px.
"""
import dataclasses

import numpy as np
import pytest

from dasp_amd import workloads as W
from oracle import centoracle as O
from tests._cases import Case, say

pytestmark = pytest.mark.gpu

# fast trig, as the benchmark runs them; fp16x1: the single-product DFT form
FORMS = {"fft32": 10, "bf16": 11, "fp16": 12, "fp16x1": 14}


def _dev():
    import torch
    return torch.device("cuda:0")


def _check_form(name, got, want, act):
    d = np.abs(got.reshape(-1, 2)[act] - want.reshape(-1, 2)[act])
    rms, mx = float(np.sqrt((d ** 2).mean())), float(d.max())
    say("  %s: max|d| = %.3g px, rms = %.3g px over %d sub-apertures" % (name, mx, rms, act.sum()))
    if name.startswith("bf16") or name.startswith("fp16x1"):
        assert rms <= 1e-3 and mx <= 1e-2
    else:
        assert mx <= 1e-4
    return mx


# ---- 80x80 n32/nfft64/nimg16 with correlation ------------------------------------------------
C5_FRAMES = 4


@pytest.fixture(scope="module")
def c5():
    import torch
    w = W.CORR80
    say("case full-size 80x80 nyquist corr: building reference, phase and oracle (%d frames)" % C5_FRAMES)
    ref = W.reference_images(w)
    c = Case(w.nsubx, w.n, w.nfft, w.nimg, obs=w.obs, correlationCentroiding=1, corrPattern=ref,
             parabolicFit=w.parabolicFit)
    phs = [w.phase(f) for f in range(C5_FRAMES)]
    want = [O.run_phase(c.params, p)[0] for p in phs]
    d_phs = torch.from_numpy(np.stack(phs)).to(_dev())
    yield c, d_phs, want
    c.free()


@pytest.mark.timeout(400)
@pytest.mark.parametrize("form", sorted(FORMS))
def test_nyquist_corr_80x80_full_size(c5, form):
    import torch
    from dasp_amd.cmod import cent
    c, d_phs, want = c5
    h = c.cc.centstruct
    cent.setForm(h, FORMS[form] % 10)
    cent.setFastTrig(h, True)
    S = c.nsubx * c.nsubx
    npup = c.nsubx * c.n
    ce = torch.full((C5_FRAMES, S, 2), -9.0, dtype=torch.float32, device=_dev())
    say("case full-size 80x80 nyquist corr %s: run_device %d frames, no image buffer" % (form, C5_FRAMES))
    cent.runDevice(h, d_phs.data_ptr(), npup * npup, C5_FRAMES, ce.data_ptr(),
                   stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    got = ce.cpu().numpy()
    act = c.cc.subflag.ravel() == 1
    for f in range(C5_FRAMES):
        _check_form("%s frame %d" % (form, f), got[f], want[f], act)
        assert np.all(got[f][~act] == 0)
    cent.setForm(h, 0)


# ---- 80x80 nfft32, the benchmark's 64-frame call ---------------------------------------------
C3_CHECK = (0, 21, 42, 63)


@pytest.fixture(scope="module")
def c3():
    import torch
    from dasp_amd import synth
    w = W.WFS80
    frames = 64
    say("case full-size 80x80 nfft32: 64 phase frames on the device, oracle on frames %s" % (C3_CHECK,))
    c = Case(w.nsubx, w.n, w.nfft, w.nimg, obs=w.obs)
    d_phs = synth.phase_screens_torch(frames, w.npup, w.D, w.r0, w.L0, w.seed, _dev(), wavelength_scale=500. / 640.)
    d_phs[0].copy_(torch.from_numpy(w.phase(0)).to(_dev()))
    torch.cuda.synchronize()
    host = {f: d_phs[f].cpu().numpy() for f in C3_CHECK}
    want = {f: O.run_phase(c.params, host[f])[0] for f in C3_CHECK}
    yield c, d_phs, want
    c.free()


@pytest.mark.timeout(300)
@pytest.mark.parametrize("form", sorted(FORMS))
def test_wfs80x80_benchmark_call_every_frame(c3, form):
    import torch
    from dasp_amd.cmod import cent
    c, d_phs, want = c3
    h = c.cc.centstruct
    cent.setForm(h, FORMS[form] % 10)
    cent.setFastTrig(h, True)
    frames = d_phs.shape[0]
    S = c.nsubx * c.nsubx
    npup = c.nsubx * c.n
    st = torch.cuda.current_stream().cuda_stream
    ce = torch.full((frames, S, 2), -9.0, dtype=torch.float32, device=_dev())
    say("case full-size 80x80 nfft32 %s: one run_device of %d frames, no image buffer" % (form, frames))
    cent.runDevice(h, d_phs.data_ptr(), npup * npup, frames, ce.data_ptr(), stream=st)
    one = torch.full((frames, S, 2), -9.0, dtype=torch.float32, device=_dev())
    for f in range(frames):   # each frame alone: the batched frame indexing adds nothing
        cent.runDevice(h, d_phs[f].data_ptr(), npup * npup, 1, one[f].data_ptr(), stream=st)
    torch.cuda.synchronize()
    got, got1 = ce.cpu().numpy(), one.cpu().numpy()
    act = c.cc.subflag.ravel() == 1
    dd = np.abs(got - got1).max()
    say("  %s: max|64-frame run - one-frame runs| = %.3g px over all frames" % (form, dd))
    assert dd <= 1e-6
    assert np.all(got[:, ~act] == 0)
    for f in C3_CHECK:
        _check_form("%s frame %d" % (form, f), got[f], want[f], act)
    cent.setForm(h, 0)


# ---- the multi-WFS set ------------------------------------------------------------------------
def _multi_geo():
    return [(W.LGS80, k) for k in range(6)] + [(W.NGS8, 6 + k) for k in range(3)] + [(W.TRUTH40, 9)]


def _multi_cases(noise):
    spots = {}
    cases = []
    for w, k in _multi_geo():
        if w.name not in spots:
            spots[w.name] = w.spots()
        kw = dict(sig=w.sig, seed=1 + k, form=10, spotpsf=spots[w.name])
        if noise:
            kw.update(readnoise=w.readnoise, addPoisson=w.addPoisson, noiseFloor=w.noiseFloor,
                      threshType=w.threshType, relativeFloor=w.relativeFloor)
        cases.append(Case(w.nsubx, w.n, w.nfft, w.nimg, obs=w.obs, **kw))
    return cases


@pytest.fixture(scope="module")
def multi_phase():
    frames = 2
    say("case full-size multi-WFS: phase of 10 WFS x %d frames" % frames)
    return [[dataclasses.replace(w, seed=10 + k).phase(f) for f in range(frames)] for w, k in _multi_geo()]


@pytest.mark.timeout(400)
@pytest.mark.parametrize("noise", [False, True], ids=["noiseless", "noise"])
def test_multi_wfs_full_set_vs_oracle(multi_phase, noise):
    import torch
    from dasp_amd.cmod import cent
    from dasp_amd.wfsbatch import WFSBatch
    frames = len(multi_phase[0])
    cases = _multi_cases(noise)
    b = WFSBatch([c.cc for c in cases], frames)
    launches, others = b.info()
    say("case full-size multi-WFS %s: launches %s, WFS outside them %d" %
        ("noise" if noise else "noiseless", [(L["np"], L["fftsize"], L["nimg"], L["spot"], L["nwfs"]) for L in launches],
         others))
    for k in range(len(cases)):
        for f in range(frames):
            b.phs[k][f].copy_(torch.from_numpy(multi_phase[k][f]))
    seed = 21
    b.draw_noise(seed=seed)
    torch.cuda.synchronize()
    b.run()
    torch.cuda.synchronize()
    worst_clean = 0.0
    for k, ((w, _), c) in enumerate(zip(_multi_geo(), cases)):
        got = b.slopes(k).reshape(frames, -1, 2)
        act = c.cc.subflag.ravel() == 1
        S = c.nsubx * c.nsubx
        m2 = c.nimg * c.nimg
        for f in range(frames):
            zp = zr = None
            if noise:   # the streams WFSBatch.draw_noise gave this WFS and frame, drawn on the host
                zp = cent.drawNormals(seed + k, 2 * f, S * m2).reshape(S, c.nimg, c.nimg)
                zr = cent.drawNormals(seed + k, 2 * f + 1, S * m2).reshape(S, c.nimg, c.nimg)
            oc, ob, _ = O.run_phase(c.params, multi_phase[k][f], zp, zr)
            d = np.abs(got[f][act] - oc[act])
            assert np.all(got[f][~act] == 0)
            if not noise:
                say("  %s #%d frame %d: max|d| = %.3g px" % (w.name, k, f, d.max()))
                assert d.max() < 1e-4
            else:
                # slopes of a sub-aperture are exact unless one of its pixels sits on a decision
                # boundary (Poisson CDF step, floor); such sub-apertures are few
                bad = (d > 1e-4).any(axis=1)
                say("  %s #%d frame %d: %d of %d sub-apertures off by > 1e-4 px" % (w.name, k, f, bad.sum(), act.sum()))
                assert bad.sum() <= max(2, 0.01 * act.sum())
                worst_clean = max(worst_clean, float(d[~bad].max()) if (~bad).any() else 0.0)
    b.free()
    for c in cases:
        c.free()
