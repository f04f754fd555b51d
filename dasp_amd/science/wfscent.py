"""The per-step logic of science/wfscent.py ``wfscent.generateNext`` (:630-733) for a phase-input
WFS, over a ``dasp_amd.util.centroid.centroid`` object.

Per step: the integration state machine (:659-673), the phase copy when the parent's array is
not the WFS's own (:680-682), ``runCalc`` (:683), and at readout the tip-tilt removal (:686-692:
the mean over all nsubx^2 entries, inactive zeros included, is subtracted from each slope axis)
and the covariance sums (:693-701). With a device-resident centroid object (``device=...``) the
phase, the slopes and the covariance sums stay in HBM; the arithmetic is the same in both routes.

The simulation framework around it (parent/child wiring, message passing, config parsing) is not
part of the hot path and is not reproduced: ``generateNext`` takes the parent's phase directly.
"""


class wfscent:
    def __init__(self, wfs, control=None):
        self.wfscentObj = wfs
        self.control = {"cal_source": 0, "zeroOutput": 0, "calcCovariance": 0, "useCmod": 1}
        if control:
            self.control.update(control)
        self.dataValid = 0
        self.imageOnly = wfs.imageOnly
        wfs.outSquare = None
        wfs.outSum = None
        wfs.outN = 0

    @property
    def outputData(self):
        return self.wfscentObj.outputData

    def generateNext(self, phs):
        """One simulation step with parent output ``phs`` (same kind of array as wfs.phase)."""
        wfs = self.wfscentObj
        self.dataValid = 0
        n = wfs.nintegrations + wfs.nlatency
        if wfs.curInteg == 0:
            additive = -1 if n == 1 else 0
        elif wfs.curInteg + 1 < wfs.nintegrations:
            additive = 1
        elif wfs.curInteg + 1 < n:
            additive = 3
        else:
            additive = 2
        if self.control["zeroOutput"]:
            wfs.curInteg = 0 if wfs.curInteg + 1 >= n else wfs.curInteg + 1
            wfs.outputData[:] = 0
            self.dataValid = 1
            return
        if phs is not wfs.phase:
            wfs.phase[:] = phs
        # runCalc advances the integration counter (the engine's on the host route, the centroid
        # object's own on the device route)
        wfs.runCalc(self.control)
        if not wfs.device:
            wfs.curInteg = 0 if wfs.curInteg + 1 >= n else wfs.curInteg + 1
        if additive in (2, -1):
            out = wfs.outputData
            if wfs.subtractTipTilt == -1 or (wfs.subtractTipTilt == 1 and self.control["cal_source"] == 0 and
                                             self.imageOnly == 0):
                N = wfs.nsubaps
                out[:, :, 0] -= out[:, :, 0].sum() / N
                out[:, :, 1] -= out[:, :, 1].sum() / N
            if self.control["calcCovariance"]:
                if wfs.outSquare is None:
                    wfs.outSquare = out ** 2
                    wfs.outSum = out.clone() if hasattr(out, "clone") else out.copy()
                    wfs.outN = 1
                else:
                    wfs.outSquare += out ** 2
                    wfs.outSum += out
                    wfs.outN += 1
            self.dataValid = 1
