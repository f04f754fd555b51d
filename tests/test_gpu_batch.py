"""GPU: the multi-WFS step (persistent work-queue launches + step graph) against per-WFS device
runs of the same handles, and against the FP64 oracle.

WFS of uneven nsubx/n/fftsize/nimg and detector options share one step; one WFS takes the general path (LGS spot
kernel) inside the same graph. The step graph is replayed with new phase written in place
between replays, each replay checked against an eager per-WFS run."""
import numpy as np
import pytest

from dasp_amd import workloads as W
from tests._cases import Case, say

pytestmark = pytest.mark.gpu

# (name, nsubx, n, nfft, nimg, extra Case kwargs)
MIX = [
    ("scao8x8_n8", 8, 8, 16, 8, {}),
    ("wfs12_n16", 12, 16, 32, 16, {}),
    ("ngs4_n32_b8", 4, 32, 64, 8, {}),
    ("wfs10_noise", 10, 16, 32, 16, dict(sig=1000.0, readnoise=0.5, addPoisson=1, noiseFloor=0.1, threshType=0, relativeFloor=True)),
    ("wfs6_n16_b4", 6, 16, 32, 8, {}),
    # brightest-pixel selection (subtracting the next value) inside the fused launch
    ("wfs8_brightest", 8, 8, 16, 8, dict(sig=100.0, useBrightest=-12)),
]


def _build(frames, with_lgs=True, fast=True):
    import torch
    from dasp_amd.cmod import cent
    from dasp_amd.wfsbatch import WFSBatch
    cases = []
    for k, (name, nsubx, n, nfft, nimg, kw) in enumerate(MIX):
        c = Case(nsubx, n, nfft, nimg, obs=0.2, seed=3 + k, form=10 if fast else 0, **kw)
        cases.append(c)
    if with_lgs:
        # spot size == fftsize: fused spot convolution; spot size 64 > fftsize: general kernels
        for P in (32, 64):
            wgeo = W.WFS("lgs", 6, 16, 32, 16, 8.0, 0.2, 0.15, 25.0, spot=True, spot_size=P, sig=1000.0)
            cases.append(Case(6, 16, 32, 16, obs=0.2, sig=1000.0, spotpsf=wgeo.spots()))
            cent.setFastTrig(cases[-1].cc.centstruct, fast)
    b = WFSBatch([c.cc for c in cases], frames)
    for k, c in enumerate(cases):
        for f in range(frames):
            w = W.WFS("b%d" % k, c.nsubx, c.n, c.nfft, c.nimg, 8.0, 0.2, 0.15, 25.0)
            b.phs[k][f].copy_(torch.from_numpy(w.phase(f)))
    b.draw_noise(seed=11)
    torch.cuda.synchronize()
    return cases, b


def _eager(cases, b):
    """Per-WFS run_device of the same handles on the same buffers' contents."""
    import torch
    from dasp_amd.cmod import cent
    out = []
    st = torch.cuda.current_stream().cuda_stream
    for k, c in enumerate(cases):
        ce = torch.full_like(b.cents[k], -9.0)
        zp = b.zpois[k].data_ptr() if b.zpois[k] is not None else 0
        zr = b.zread[k].data_ptr() if b.zread[k] is not None else 0
        npup = c.nsubx * c.n
        cent.runDevice(c.cc.centstruct, b.phs[k].data_ptr(), npup * npup, b.nframes, ce.data_ptr(),
                       zpois_ptr=zp, zread_ptr=zr, stream=st)
        out.append(ce)
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in out]


def _compare(cases, b, want, tag):
    worst = 0.0
    for k, c in enumerate(cases):
        got = b.slopes(k)
        d = np.abs(got - want[k]).max()
        worst = max(worst, d)
        say("%s %s: max|batch - per-WFS| = %.3g px" % (tag, MIX[k][0] if k < len(MIX) else "lgs%d" % k, d))
        # the batch kernel and the single-WFS kernel are separate compilations of the same item
        # code: the compiler may contract a different multiply-add, a rounding-level difference
        assert d <= 1e-5
    return worst


def test_batch_matches_per_wfs_and_replays():
    import torch
    say("case batch mix, 3 frames, graph replays")
    cases, b = _build(3)
    launches, others = b.info()
    say("launches %s, others %d" % (launches, others))
    assert others == 1                       # the 64-pixel-spot LGS runs its own kernels in the graph
    assert sorted((L["np"], L["fftsize"], L["nimg"], L["spot"]) for L in launches) == \
        sorted({(8, 16, 8, False), (16, 32, 16, False), (32, 64, 8, False), (16, 32, 8, False), (16, 32, 16, True)})
    for rep in range(3):
        for k in range(len(cases)):          # new phase in place, eager work between replays
            b.phs[k].mul_(-1.0 if rep % 2 else 1.25)
        torch.cuda.synchronize()
        for ce in b.cents:
            ce.fill_(-7.0)
        b.run()
        torch.cuda.synchronize()
        want = _eager(cases, b)
        _compare(cases, b, want, "replay %d" % rep)
    b.free()
    for c in cases:
        c.free()


def test_batch_vs_oracle_and_inactive_zero():
    """Noiseless members of the step against the FP64 oracle (1e-4 px); inactive slopes zero."""
    import torch
    from oracle import centoracle as O
    say("case batch vs oracle")
    cases, b = _build(2, with_lgs=False)
    for ce in b.cents:
        ce.fill_(-7.0)
    b.run()
    torch.cuda.synchronize()
    for k, c in enumerate(cases):
        if MIX[k][5]:
            continue                         # noisy member: covered by the per-WFS comparison
        got = b.slopes(k)
        act = c.cc.subflag.ravel() == 1
        for f in range(b.nframes):
            phs = b.phs[k][f].cpu().numpy()
            oc, _, _ = O.run_phase(c.params, phs, None, None)
            d = np.abs(got[f].reshape(-1, 2) - oc.reshape(-1, 2))[act]
            say("%s frame %d: max|d| = %.3g px" % (MIX[k][0], f, d.max()))
            assert d.max() < 1e-4
            assert np.all(got[f].reshape(-1, 2)[~act] == 0)
    b.free()
    for c in cases:
        c.free()


def test_batch_tune_and_update_rebuild():
    """The autotuner keeps results unchanged; an update (REFCENTS) rebuilds the step graph."""
    import torch
    say("case batch tune + update")
    cases, b = _build(2, with_lgs=False)
    b.run()
    torch.cuda.synchronize()
    before = [b.slopes(k) for k in range(len(cases))]
    launches, _ = b.tune(reps=2)
    say("tuned: %s" % launches)
    assert all(L["tuned_ms"] > 0 and L["workgroups"] > 0 for L in launches)
    b.run()
    torch.cuda.synchronize()
    for k in range(len(cases)):
        np.testing.assert_array_equal(b.slopes(k), before[k])
    ref = np.full((cases[0].nsubx, cases[0].nsubx, 2), 0.25, np.float32)
    cases[0].cc.update(18, ref)              # REFCENTS on the first WFS only
    b.run()
    torch.cuda.synchronize()
    got = b.slopes(0)
    act = cases[0].cc.subflag == 1
    np.testing.assert_allclose(got[:, act], before[0][:, act] - 0.25, rtol=0, atol=1e-6)
    np.testing.assert_array_equal(b.slopes(1), before[1])
    b.free()
    for c in cases:
        c.free()


def test_batch_with_dft_form_members():
    """WFS on the DFT-GEMM forms (n16/nfft32 FP16, n32/nfft64 BF16 with the correlation stage) run
    their own kernels inside the step graph beside a fused member: each equals its eager per-WFS
    device run, over graph replays with new phase written in place."""
    import torch
    from dasp_amd.cmod import cent
    from dasp_amd.wfsbatch import WFSBatch
    say("case batch with DFT-form members")
    w16 = W.WFS("d16", 11, 16, 32, 16, 8.0, 0.2, 0.15, 25.0)
    w32 = W.WFS("d32", 8, 32, 64, 16, 8.0, 0.2, 0.15, 25.0)
    cases = [Case(11, 16, 32, 16, obs=0.2, form=12),
             Case(8, 32, 64, 16, obs=0.2, form=11, correlationCentroiding=1,
                  corrPattern=W.reference_images(w32), parabolicFit=5),
             Case(12, 16, 32, 16, obs=0.2, form=10)]
    frames = 2
    b = WFSBatch([c.cc for c in cases], frames)
    launches, others = b.info()
    say("launches %s, others %d" % (launches, others))
    assert others == 2
    wgeo = [w16, w32, W.WFS("f16", 12, 16, 32, 16, 8.0, 0.2, 0.15, 25.0)]
    for k in range(len(cases)):
        for f in range(frames):
            b.phs[k][f].copy_(torch.from_numpy(wgeo[k].phase(f)))
    for rep in range(2):
        for k in range(len(cases)):
            b.phs[k].mul_(1.0 if rep == 0 else -0.8)
        torch.cuda.synchronize()
        for ce in b.cents:
            ce.fill_(-7.0)
        b.run()
        torch.cuda.synchronize()
        st = torch.cuda.current_stream().cuda_stream
        for k, c in enumerate(cases):
            ce = torch.full_like(b.cents[k], -9.0)
            npup = c.nsubx * c.n
            nact, _, csz = cent.info(c.cc.centstruct)
            bi = torch.zeros((frames, c.nsubx * c.nsubx, c.nimg, c.nimg), dtype=torch.float32, device="cuda:0")
            ci = torch.zeros((frames, c.nsubx * c.nsubx, csz, csz), dtype=torch.float32, device="cuda:0")
            cent.runDevice(c.cc.centstruct, b.phs[k].data_ptr(), npup * npup, frames, ce.data_ptr(), bi.data_ptr(),
                           ci.data_ptr(), stream=st)
            torch.cuda.synchronize()
            got, want = b.slopes(k), ce.cpu().numpy()
            d = np.abs(got - want).max()
            say("replay %d member %d: max|batch - per-WFS| = %.3g px" % (rep, k, d))
            assert d <= 1e-5
    b.free()
    for c in cases:
        c.free()


def test_batch_rebuilds_after_member_scratch_grows():
    """A member that runs its own kernels in the step graph (n32 DFT form with the correlation
    stage, no image buffer given) uses the handle's image scratch. A device run of that handle
    with more frames moves the scratch; the next step must rebuild its graph on the new address
    (the old one is freed) and still equal the eager per-WFS run, on two steps in a row."""
    import torch
    from dasp_amd.cmod import cent
    from dasp_amd.wfsbatch import WFSBatch
    say("case batch member scratch grows between steps")
    w32 = W.WFS("d32", 8, 32, 64, 16, 8.0, 0.2, 0.15, 25.0)
    cases = [Case(8, 32, 64, 16, obs=0.2, form=11, correlationCentroiding=1,
                  corrPattern=W.reference_images(w32), parabolicFit=5),
             Case(12, 16, 32, 16, obs=0.2, form=10)]
    frames = 2
    b = WFSBatch([c.cc for c in cases], frames)
    wgeo = [w32, W.WFS("f16", 12, 16, 32, 16, 8.0, 0.2, 0.15, 25.0)]
    for k in range(len(cases)):
        for f in range(frames):
            b.phs[k][f].copy_(torch.from_numpy(wgeo[k].phase(f)))
    torch.cuda.synchronize()
    b.run()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream().cuda_stream
    c0 = cases[0]
    npup = c0.nsubx * c0.n
    big = 5   # more frames than the batch: the handle's image scratch is reallocated
    ph = torch.zeros((big, npup, npup), dtype=torch.float32, device="cuda:0")
    ph[:frames].copy_(b.phs[0])
    ph[frames:].copy_(b.phs[0][:1])
    ce_big = torch.zeros((big, c0.nsubx * c0.nsubx, 2), dtype=torch.float32, device="cuda:0")
    cent.runDevice(c0.cc.centstruct, ph.data_ptr(), npup * npup, big, ce_big.data_ptr(), stream=st)
    torch.cuda.synchronize()
    for rep in range(2):
        for ce in b.cents:
            ce.fill_(-7.0)
        b.run()
        torch.cuda.synchronize()
        for k, c in enumerate(cases):
            ce = torch.full_like(b.cents[k], -9.0)
            n2 = (c.nsubx * c.n) ** 2
            nact, _, csz = cent.info(c.cc.centstruct)
            bi = torch.zeros((frames, c.nsubx * c.nsubx, c.nimg, c.nimg), dtype=torch.float32, device="cuda:0")
            ci = torch.zeros((frames, c.nsubx * c.nsubx, csz, csz), dtype=torch.float32, device="cuda:0")
            cent.runDevice(c.cc.centstruct, b.phs[k].data_ptr(), n2, frames, ce.data_ptr(), bi.data_ptr(),
                           ci.data_ptr(), stream=st)
            torch.cuda.synchronize()
            d = np.abs(b.slopes(k) - ce.cpu().numpy()).max()
            say("step %d member %d after scratch growth: max|batch - per-WFS| = %.3g px" % (rep, k, d))
            assert d <= 1e-5
    # the big run's first frames equal the batch's frames
    d = np.abs(ce_big[:frames].cpu().numpy().reshape(frames, -1) - b.slopes(0).reshape(frames, -1)).max()
    assert d <= 1e-5
    b.free()
    for c in cases:
        c.free()
