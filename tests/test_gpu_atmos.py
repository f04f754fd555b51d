"""GPU: the phase producer upstream of the WFS on the device (include/dasp_atmos.h) against the
oracle of oracle/atmosoracle.py: createPupilPhs for a natural guide star, makeLayer on ring-buffered
screens, the DM's spline surface and line of sight, and a closed loop (atmosphere -> mirror -> WFS
-> reconstructor -> mirror) that never leaves HBM, against the same loop composed on the host."""
import numpy as np
import pytest

from oracle import atmosoracle as A
from tests._cases import say

pytestmark = pytest.mark.gpu


def _dev():
    import torch
    return torch.device("cuda:0")


def _pupil(npup, obs=0.2):
    yy, xx = np.mgrid[0:npup, 0:npup] - (npup - 1) / 2.
    r = np.hypot(yy, xx)
    return ((r <= npup / 2.) & (r >= npup / 2. * obs)).astype(np.float32)


def _layers(rng, npup, specs):
    out = []
    for k, (rows, cols, x0, y0, pc, pr, ca, ra) in enumerate(specs):
        out.append(dict(key="L%d" % k, screen=rng.standard_normal((rows, cols)) * (k + 1), x0=x0, y0=y0, pos_col=pc,
                        pos_row=pr, col_add=ca, row_add=ra))
    return out


def _oracle_layers(layers, icol, irow):
    return [dict(key=L["key"], x0=L["x0"], y0=L["y0"], pos_col=L["pos_col"], pos_row=L["pos_row"],
                 col_add=L["col_add"], row_add=L["row_add"], interp_col=icol[L["key"]], interp_row=irow[L["key"]])
            for L in layers]


SPECS = [(80, 84, 4, 3, 0.0, 0.0, -1, -1), (78, 90, 7, 5, 0.35, 0.8, 1, -1), (90, 76, 2, 9, 0.6, 0.15, -1, 1)]


@pytest.mark.parametrize("full_pupil,lam", [(False, 640.), (True, 500.)])
def test_pupil_phase_vs_oracle(full_pupil, lam):
    import torch
    from dasp_amd.science.atmos import DeviceAtmos
    npup = 64
    say("case atmos pupil phase full_pupil=%s lam=%g" % (full_pupil, lam))
    rng = np.random.default_rng(1)
    layers = _layers(rng, npup, SPECS)
    pup = None if full_pupil else _pupil(npup)
    da = DeviceAtmos(npup, layers, pupfn=pup, source_lam=lam)
    icol = {"L0": 0.25, "L1": 0.0, "L2": 0.7}
    irow = {"L0": 0.5, "L1": 0.33, "L2": 0.0}
    o64 = torch.empty((npup, npup), dtype=torch.float64, device=_dev())
    o32 = torch.empty((npup, npup), dtype=torch.float32, device=_dev())
    da.create_pupil_phs(icol, irow, out64=o64, out32=o32)
    torch.cuda.synchronize()
    want = A.create_pupil_phs({L["key"]: L["screen"] for L in layers}, _oracle_layers(layers, icol, irow), npup,
                              pupfn=pup, remove_piston=True, source_lam=lam)
    got = o64.cpu().numpy()
    err = np.abs(got - want).max()
    say("  max |device - oracle| = %.3g (phase rms %.3g)" % (err, want.std()))
    assert err < 1e-12 * np.abs(want).max()
    assert np.array_equal(o32.cpu().numpy(), got.astype(np.float32))
    if pup is not None:
        assert np.all(got[pup == 0] == 0)


def test_layer_ring_updates_vs_oracle():
    """Five steps of makeLayer (new columns and rows on every side) and createPupilPhs: the ring
    buffers give the shifted screens of the reference."""
    import torch
    from dasp_amd.science.atmos import DeviceAtmos
    npup = 48
    say("case atmos ring updates")
    rng = np.random.default_rng(2)
    specs = [(60, 64, 5, 4, 0.0, 0.0, -1, -1), (58, 62, 3, 2, 0.2, 0.4, 1, 1), (62, 60, 6, 7, 0.5, 0.1, -1, 1)]
    layers = _layers(rng, npup, specs)
    pup = _pupil(npup)
    da = DeviceAtmos(npup, layers, pupfn=pup, source_lam=700.)
    screens = {L["key"]: L["screen"].copy() for L in layers}
    o64 = torch.empty((npup, npup), dtype=torch.float64, device=_dev())
    for step in range(5):
        icol, irow = {}, {}
        for L in layers:
            H, Wd = screens[L["key"]].shape
            nc, nr = 1 + (step % 2), step % 3
            cols = rng.standard_normal((nc, H))
            rows = rng.standard_normal((nr, Wd))
            screens[L["key"]] = A.make_layer(screens[L["key"]], cols, rows, L["col_add"], L["row_add"])
            da.make_layer(L["key"], cols, rows)
            icol[L["key"]] = float(rng.uniform(0, 1))
            irow[L["key"]] = float(rng.uniform(0, 1)) if step % 2 else 0.0
        da.create_pupil_phs(icol, irow, out64=o64)
        torch.cuda.synchronize()
        for L in layers:
            assert np.array_equal(da.screen(L["key"]).cpu().numpy(), screens[L["key"]]), (step, L["key"])
        want = A.create_pupil_phs(screens, _oracle_layers(layers, icol, irow), npup, pupfn=pup, source_lam=700.)
        err = np.abs(o64.cpu().numpy() - want).max()
        say("  step %d: max |device - oracle| = %.3g" % (step, err))
        assert err < 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("npup,nact,coupling", [(33, 9, None), (64, 17, 0.1), (97, 13, 0.25)])
def test_dm_surface_vs_oracle(npup, nact, coupling):
    import torch
    from dasp_amd.science.dm import DeviceMirrorSurface
    say("case DM surface npup %d nact %d coupling %s" % (npup, nact, coupling))
    rng = np.random.default_rng(npup)
    act = rng.standard_normal((nact, nact)).astype(np.float32)
    ms = DeviceMirrorSurface(npup, nact, actCoupling=coupling)
    got = ms.fit(torch.from_numpy(act).to(_dev())).cpu().numpy()
    want = A.mirror_fit_spline(act, npup, act_coupling=coupling)
    err = np.abs(got - want).max()
    say("  max |device - oracle| = %.3g of max %.3g" % (err, np.abs(want).max()))
    # double-precision sums in another order, then the float32 cast: a few ulp at most
    assert err <= 4 * np.spacing(np.float32(np.abs(want).max()))


@pytest.mark.parametrize("base,conj,wla", [("f64", 0., 1.), ("f32", -1., 1.), ("none", 0., 0.8), ("f64", -2., 0.8)])
def test_dm_line_of_sight_vs_oracle(base, conj, wla):
    import torch
    from dasp_amd.science.dm import DeviceDMLineOfSight
    npup, dmpup, yoff, xoff = 56, 64, 3, 5
    say("case DM line of sight base=%s conj=%g wla=%g" % (base, conj, wla))
    rng = np.random.default_rng(4)
    pup = _pupil(npup)
    atm64 = rng.standard_normal((npup, npup))
    surf = rng.standard_normal((dmpup, dmpup)).astype(np.float32)
    los = DeviceDMLineOfSight(dmpup, npup, pup, yoff, xoff, conjHeight=conj, wavelengthAdjustor=wla)
    out = torch.zeros((npup, npup), dtype=torch.float32, device=_dev())
    d_surf = torch.from_numpy(surf).to(_dev())
    if base == "f64":
        atm = torch.from_numpy(atm64).to(_dev())
        start = atm64.astype(np.float32)
    elif base == "f32":
        out.copy_(torch.from_numpy(atm64.astype(np.float32)))
        atm = out                                    # in place, as xinterp_dm's outputData
        start = atm64.astype(np.float32)
    else:
        atm = None
        start = np.zeros((npup, npup), np.float32)
    los.selectSubPupil(out, d_surf, atmos=atm)
    torch.cuda.synchronize()
    want = A.select_sub_pupil(start, surf, yoff, xoff, pupfn=pup, add_to_output=atm is not None, conj_height=conj,
                              wavelength_adjustor=wla)
    got = out.cpu().numpy()
    err = np.abs(got - want).max()
    say("  max |device - oracle| = %.3g" % err)
    assert err <= 2e-6 * np.abs(want).max()
    assert np.all(got[pup == 0] == 0)


def test_closed_loop_in_hbm():
    """Three steps of atmosphere -> ground-conjugate spline mirror -> 8x8 WFS (device route, the
    mirror writing straight into the WFS's phase buffer) -> reconstructor -> next step's mirror,
    everything in HBM; the same loop on the host from the oracles (atmosphere, mirror, centroids)
    and NumPy (reconstructor) gives the same slopes and actuators."""
    import torch
    from dasp_amd.cmod import cent
    from dasp_amd.science.atmos import DeviceAtmos
    from dasp_amd.science.dm import DeviceXinterpDM
    from dasp_amd.science.tomorecon import DeviceRecon, cent_index
    from oracle import centoracle as O
    from tests._cases import Case
    say("case closed loop atmosphere -> DM -> WFS -> recon in HBM")
    nsubx, n, nfft, nimg = 8, 8, 16, 8
    npup = nsubx * n
    nact = nsubx + 1
    rng = np.random.default_rng(8)
    c = Case(nsubx, n, nfft, nimg, obs=0.25)
    pup = c.cc.pupfn.transpose(0, 2, 1, 3).reshape(npup, npup).astype(np.float32)   # per-subaperture -> 2-D
    layers = _layers(rng, npup, [(80, 80, 4, 4, 0.0, 0.0, -1, -1), (76, 82, 5, 2, 0.3, 0.6, 1, -1)])
    for L in layers:
        L["screen"] = L["screen"] * 0.4
    da = DeviceAtmos(npup, layers, pupfn=pup, source_lam=640.)
    dm = DeviceXinterpDM(npup, nact, pup, actCoupling=0.1)
    flags = [c.cc.subflag]
    ncents = 2 * int(c.cc.subflag.sum())
    R = (rng.standard_normal((ncents, nact * nact)) * 0.05).astype(np.float32)
    rec = DeviceRecon(flags, torch.from_numpy(R).to(_dev()), gains=0.3, decay=0.99)
    phs32 = torch.zeros((npup, npup), dtype=torch.float32, device=_dev())
    atm64 = torch.empty((npup, npup), dtype=torch.float64, device=_dev())
    st = torch.cuda.current_stream().cuda_stream
    screens = {L["key"]: L["screen"].copy() for L in layers}
    acts_h = np.zeros(nact * nact)
    index = cent_index(flags, rec.offsets)
    for step in range(3):
        icol = {L["key"]: 0.1 + 0.3 * step for L in layers}
        irow = {L["key"]: 0.2 * step for L in layers}
        # device: atmosphere, mirror (last step's actuators), WFS into the recon's slope buffer, recon
        da.create_pupil_phs(icol, irow, out64=atm64)
        dm.generateNext(phs32, atmos=atm64, actmap=rec.acts[0].view(nact, nact))
        cent.runDevice(c.cc.centstruct, phs32.data_ptr(), npup * npup, 1, rec.slope_buffer(0).data_ptr(), stream=st)
        rec.run()
        torch.cuda.synchronize()
        # host: the same loop from the oracles
        atm = A.create_pupil_phs(screens, _oracle_layers(layers, icol, irow), npup, pupfn=pup, source_lam=640.)
        surf = A.mirror_fit_spline(acts_h.astype(np.float32).reshape(nact, nact), npup, act_coupling=0.1)
        ph = A.select_sub_pupil(atm.astype(np.float32), surf, 0, 0, pupfn=pup)
        assert np.abs(phs32.cpu().numpy() - ph).max() <= 2e-6 * np.abs(ph).max()
        oc, _, _ = O.run_phase(c.params, ph)
        slopes_h = np.zeros((nsubx * nsubx, 2))
        act = c.cc.subflag.ravel() == 1
        slopes_h[act] = oc[act]
        got_s = rec.slope_buffer(0).cpu().numpy().reshape(-1, 2)
        ds = np.abs(got_s - slopes_h).max()
        data = slopes_h.ravel()[index]
        acts_h = 0.99 * acts_h - 0.3 * (data @ R.astype(np.float64))
        da_ = np.abs(rec.acts[0].cpu().numpy() - acts_h).max()
        say("  step %d: slopes max|d| %.3g px, actuators max|d| %.3g (max %.3g)" % (step, ds, da_, np.abs(acts_h).max()))
        assert ds < 1e-4
        assert da_ < 1e-4 * max(1.0, np.abs(acts_h).max())
        acts_h = rec.acts[0].cpu().numpy().astype(np.float64)   # carry the device's actuators (float32)
    rec.free()
    c.free()
