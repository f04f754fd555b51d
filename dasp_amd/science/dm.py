"""The deformable mirror on the device: util/dm.py ``MirrorSurface`` of type "spline" (:1268-1536,
:1723-1740) and ``DMLineOfSight.selectSubPupil`` (:2057-2151) for a ground-conjugate mirror, the
two steps of science/xinterp_dm.py ``generateNext`` (:407-513) between the reconstructor's
actuator commands and the WFS's phase (include/dasp_atmos.h, dasp_amd/csrc/atmos.hip).

MirrorSurface.fit interpolates the coupled actuator map with gslCubSplineIntrp's two natural
cubic spline passes. A natural spline through fixed knots is linear in the knot values, so both
passes are one fixed (npup x nact) operator B (the spline of each unit vector at every pixel,
built once on the host in double): surface = B F B^T, two small products on the device. The
actuator map stays in HBM from the reconstructor to the mirror, the surface stays for the line of
sight, and the line of sight writes straight into the WFS's phase buffer, so nothing is copied
to the host (science/wfscent.py:680-682 finds the phase already in place).
"""
import numpy as np

from dasp_amd.cmod import cent


class DeviceMirrorSurface:
    """MirrorSurface(typ="spline") on an npup x npup surface from an nact x nact actuator map.
    actCoupling: MirrorSurface.fudge's coupling (None or 0: none); actoffset 0 (Fried geometry:
    actuators at the corners of the sub-apertures; the spline is not defined beyond the outer
    actuators, as GSL refuses such pixels)."""

    def __init__(self, npup, nact, actCoupling=None, actoffset=0., device="cuda:0"):
        import torch
        self.torch = torch
        self.dev = torch.device(device)
        self.npup, self.nact = int(npup), int(nact)
        self.coupling = float(actCoupling or 0.)
        op = cent.dmSplineOperator(self.npup, self.nact, float(actoffset))
        self.op = torch.from_numpy(np.ascontiguousarray(op)).to(self.dev)
        self.opT = torch.from_numpy(np.ascontiguousarray(op.T)).to(self.dev)
        self.phsOut = torch.zeros((self.npup, self.npup), dtype=torch.float32, device=self.dev)
        nbytes = ((self.nact * self.nact * 4 + 7) // 8) * 8 + self.npup * self.nact * 8
        self.work = torch.empty(nbytes // 8 + 1, dtype=torch.float64, device=self.dev)

    def fit(self, actmap, phsOut=None, stream=None):
        """The surface of the actuator map (float32 (nact, nact) device tensor) into phsOut."""
        torch = self.torch
        out = self.phsOut if phsOut is None else phsOut
        a = actmap.contiguous()
        if a.dtype != torch.float32 or tuple(a.shape) != (self.nact, self.nact):
            raise ValueError("actmap must be float32 (%d, %d)" % (self.nact, self.nact))
        st = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
        cent.dmSurface(a.data_ptr(), self.nact, self.coupling, self.op.data_ptr(), self.opT.data_ptr(), self.npup,
                       out.data_ptr(), self.work.data_ptr(), stream=st)
        return out


class DeviceDMLineOfSight:
    """DMLineOfSight for a ground-conjugate mirror without sub-pixel interpolation or alignment
    offset: the selected window (yoff, xoff) of the dmpup x dmpup surface, the wavelength adjustor,
    the sign of the conjugate height, then the pupil and the piston inside it."""

    def __init__(self, dmpup, npup, pupfn, yoff=0, xoff=0, conjHeight=0., wavelengthAdjustor=1., device="cuda:0"):
        import torch
        self.torch = torch
        self.dev = torch.device(device)
        self.dmpup, self.npup = int(dmpup), int(npup)
        self.yoff, self.xoff = int(yoff), int(xoff)
        self.conjHeight = float(conjHeight)
        self.wavelengthAdjustor = float(wavelengthAdjustor)
        if pupfn is None:
            self.pupfn, self.pupsum = None, 0.
        else:
            p = np.ascontiguousarray(pupfn, np.float32)
            self.pupfn = torch.from_numpy(p).to(self.dev)
            self.pupsum = float(p.astype(np.float64).sum())
        self.work = torch.empty(cent.atmosWorkDoubles(self.npup), dtype=torch.float64, device=self.dev)

    def selectSubPupil(self, outputData, surface, atmos=None, removePiston=1, stream=None):
        """outputData (float32 (npup, npup) device tensor) = atmosphere + the mirror's window, pupil,
        piston removed. atmos: the atmosphere's phase (float64 or float32 device tensor, may be
        outputData itself), or None for a mirror without atmosphere parent."""
        torch = self.torch
        wla = self.wavelengthAdjustor
        adjust = wla != 1.
        factor = -wla if self.conjHeight < 0 else wla
        b64 = atmos.data_ptr() if atmos is not None and atmos.dtype == torch.float64 else 0
        b32 = atmos.data_ptr() if atmos is not None and atmos.dtype == torch.float32 else 0
        st = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
        cent.dmLineOfSight(surface.data_ptr(), self.dmpup, self.yoff, self.xoff, self.npup, factor, adjust, b64, b32,
                           self.pupfn.data_ptr() if self.pupfn is not None else 0, self.pupsum,
                           bool(removePiston) and self.pupfn is not None, outputData.data_ptr(), self.work.data_ptr(),
                           stream=st)
        return outputData


class DeviceXinterpDM:
    """science/xinterp_dm.py generateNext for a ground-conjugate spline mirror: with new actuator
    commands (dm_update) the surface is refitted; then the line of sight adds it to the
    atmosphere's phase in the WFS's float32 phase buffer."""

    def __init__(self, npup, nact, pupfn, actCoupling=None, dmpup=None, yoff=0, xoff=0, conjHeight=0.,
                 wavelengthAdjustor=1., device="cuda:0"):
        dmpup = dmpup or npup
        self.surface = DeviceMirrorSurface(dmpup, nact, actCoupling, device=device)
        self.los = DeviceDMLineOfSight(dmpup, npup, pupfn, yoff, xoff, conjHeight, wavelengthAdjustor, device=device)

    def generateNext(self, outputData, atmos=None, actmap=None, stream=None):
        if actmap is not None:
            self.surface.fit(actmap, stream=stream)
        return self.los.selectSubPupil(outputData, self.surface.phsOut, atmos=atmos, stream=stream)
