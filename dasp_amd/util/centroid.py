"""The hot-path part of util/centroid.py's ``centroid`` class over the gfx950 engine.

Covers what sits on the phase -> slopes path: the derived sizes and integration counts of the
constructor (util/centroid.py:115-340), ``initMem`` (:442-493, output and image buffers),
``initialiseCmod`` (:495-520), the ``runCalc`` / ``runCmod`` / ``runSlopeCalc`` dispatch
(:547-615), ``takeReference`` / ``setRefSlopes`` (:1536-1572) and ``takeCorrImage`` (:1575-1633).
Python-side linearisation (``linearSteps``) and the pure-Python imaging route are not part of
it: calibration runs in the engine through ``update(CALDATA, ...)``.

``device=None`` keeps the reference's host buffers (numpy, the engine copies phase in and
slopes out every call). ``device="cuda:0"`` keeps ``phase``, ``outputData`` and the images as
torch tensors in HBM: ``runCalc`` then calls ``cent.runDevice`` on them, draws the readout's
noise with the engine's host generator (same streams as the host route, so both routes give
the same numbers) and runs the integration cycle itself.
"""
import numpy

from dasp_amd.cmod import cent
from dasp_amd.util import centcmod as centcmod_mod
from dasp_amd.util import correlation


class centroid:
    def __init__(self, nsubx, pup=None, phase=None, readnoise=0., readbg=0., addPoisson=0, noiseFloor=0.,
                 sig=1., skybrightness=0., phasesize=None, fftsize=None, clipsize=None, nimg=None, ncen=None,
                 tstep=0.05, integtime=0.05, latency=0., wfs_minarea=0.5, spotpsf=None, centroidPower=None,
                 opticalBinning=0, subtractTipTilt=0, phaseMultiplier=1, centWeight=None,
                 correlationCentroiding=0, corrThresh=0., corrPattern=None, threshType=0, imageOnly=0,
                 useBrightest=0, preBinningFactor=1, parabolicFit=0, gaussianFitVals=None, seed=1,
                 device=None):
        self.nsubx = nsubx
        self.pup = pup
        self.phasesize = phasesize
        self.fftsize = fftsize
        self.clipsize = clipsize if clipsize is not None else fftsize
        self.nimg = nimg if nimg is not None else self.clipsize
        self.ncen = ncen if ncen is not None else self.nimg
        self.readnoise = readnoise
        self.readbg = readbg
        self.addPoisson = addPoisson
        self.noiseFloor = noiseFloor
        self.sig = sig
        self.skybrightness = skybrightness
        self.tstep = tstep
        self.integtime = integtime
        self.latency = latency
        # util/centroid.py:253-259
        self.nintegrations = int(numpy.ceil(self.integtime / self.tstep))
        self.nlatency = int(numpy.ceil(self.latency / self.tstep))
        self.curInteg = 0
        self.wfs_minarea = wfs_minarea
        self.psf = spotpsf
        self.centroidPower = centroidPower
        self.opticalBinning = opticalBinning
        self.subtractTipTilt = subtractTipTilt
        self.phaseMultiplier = phaseMultiplier
        self.centWeight = centWeight
        self.correlationCentroiding = correlationCentroiding or 0
        self.corrThresh = corrThresh
        self.corrPatternUser = corrPattern
        self.threshType = threshType
        self.imageOnly = imageOnly
        self.useBrightest = useBrightest
        self.preBinningFactor = preBinningFactor
        self.parabolicFit = parabolicFit
        self.gaussianFitVals = gaussianFitVals
        self.seed = seed
        self.linearSteps = None
        self.calNCoeff = 0
        self.magicCentroiding = 0
        self.inputImage = None
        self.refCents = None
        self.device = device
        # correlation pattern (util/centroid.py:296-306). The engine fixes the pattern size at
        # initialise (cmod/src/centmodule.c:3000-3007); without a pattern yet, an all-zero one of
        # nimg x nimg holds that size until takeCorrImage measures the real one.
        self._engine_corr = None
        if self.correlationCentroiding:
            if corrPattern is None:
                self.corrPattern = None
                shape = (nsubx, nsubx, self.nimg, self.nimg)
                self._engine_corr = numpy.zeros(shape, numpy.float32)
            else:
                if self.correlationCentroiding == 1:
                    self.corrPattern = correlation.transformPSF(self.corrPatternUser)
                else:
                    self.corrPattern = self.corrPatternUser.copy()
                self._engine_corr = self.corrPattern
                shape = corrPattern.shape
            self.corrimg = numpy.zeros(shape, numpy.float32)
        else:
            self.corrPattern = None
            self.corrimg = None
        npup = nsubx * phasesize
        if phase is None:
            phase = numpy.zeros((npup, npup), numpy.float32)
        self._host_phase = numpy.ascontiguousarray(phase, numpy.float32) if not self._is_tensor(phase) \
            else numpy.zeros((npup, npup), numpy.float32)
        self.phase = phase
        self.centcmod = None
        self._readouts = 0

    @staticmethod
    def _is_tensor(x):
        return type(x).__module__.startswith("torch")

    # ---- util/centroid.py:442-493 ------------------------------------------------------------
    def initMem(self):
        nsubx, nimg = self.nsubx, self.nimg
        if self.imageOnly == 0:
            shape = (nsubx, nsubx, 2)
        elif self.imageOnly == 1:
            shape = (nsubx, nsubx, nimg, nimg)
        else:
            shape = (nsubx * nimg, nsubx * nimg)
        # host buffers the engine writes on the host route (cmodbimg is float32, as the
        # reference's view of its float64 bimg storage)
        self._host_out = numpy.zeros((nsubx, nsubx, 2), numpy.float32)
        self._host_bimg = numpy.zeros((nsubx, nsubx, nimg, nimg), numpy.float32)
        if self.device is None:
            self.outputData = numpy.zeros(shape, numpy.float32) if self.imageOnly else self._host_out
            self.cmodbimg = self._host_bimg
            if self._is_tensor(self.phase):
                raise ValueError("a torch phase needs device=...")
        else:
            import torch
            dev = torch.device(self.device)
            self._torch = torch
            if not self._is_tensor(self.phase):
                self.phase = torch.from_numpy(numpy.ascontiguousarray(self.phase, numpy.float32)).to(dev)
            self.outputData = torch.zeros(shape, dtype=torch.float32, device=dev)
            self._dev_cents = self.outputData if self.imageOnly == 0 else \
                torch.zeros((nsubx, nsubx, 2), dtype=torch.float32, device=dev)
            self.cmodbimg = torch.zeros((nsubx, nsubx, nimg, nimg), dtype=torch.float32, device=dev)
            self._dev_corrimg = None
            if self.corrimg is not None:
                self._dev_corrimg = torch.zeros(self.corrimg.shape, dtype=torch.float32, device=dev)
            S = nsubx * nsubx
            self._zp = torch.zeros((S, nimg, nimg), dtype=torch.float32, device=dev)
            self._zr = torch.zeros((S, nimg, nimg), dtype=torch.float32, device=dev)
        self.centx = self.outputData[:, :, 0] if self.imageOnly == 0 else None
        self.centy = self.outputData[:, :, 1] if self.imageOnly == 0 else None

    # ---- util/centroid.py:495-520 ------------------------------------------------------------
    def initialiseCmod(self, nthreads=8, calsource=0, seed=None):
        self.nthreads = nthreads
        sig = self.sig.ravel() if isinstance(self.sig, numpy.ndarray) else self.sig
        self.centcmod = centcmod_mod.centcmod(
            nthreads, self.nsubx, self.ncen, self.fftsize, self.clipsize, self.nimg, self.phasesize,
            self.readnoise, self.readbg, self.addPoisson, self.noiseFloor, sig, self.skybrightness, calsource,
            self.centroidPower, self.nintegrations, self.nlatency, self.seed if seed is None else seed,
            self._host_phase, self.pup, self.psf, self._host_out, self._host_bimg, self.wfs_minarea,
            self.opticalBinning, self.centWeight, self.correlationCentroiding, self.corrThresh,
            self._engine_corr, self.corrimg, self.threshType, self.imageOnly, self.useBrightest,
            self.preBinningFactor, self.parabolicFit, self.gaussianFitVals)
        self._calsource = calsource

    # ---- util/centroid.py:547-615 ------------------------------------------------------------
    def runSlopeCalc(self, control):
        raise NotImplementedError("image-input WFS: construct centcmod with inputImage and call runSlope")

    def runCalc(self, control={"cal_source": 0}):
        if self.phaseMultiplier != 1:
            self.phase *= self.phaseMultiplier       # in place, as the reference
        self.runCmod(control["cal_source"])
        # refCents are subtracted by the engine (no Python-side linearisation here)

    def runCmod(self, calsource):
        if self.device is None:
            if self._is_tensor(self.phase):
                raise ValueError("host route needs a numpy phase")
            if self.phase is not self._host_phase:
                self._host_phase[:] = self.phase
            self.centcmod.run(calsource)
            if self.imageOnly == 1:
                self.outputData[:] = self.cmodbimg
            elif self.imageOnly == 2:
                for i in range(self.nsubx):
                    for j in range(self.nsubx):
                        self.outputData[i * self.nimg:(i + 1) * self.nimg, j * self.nimg:(j + 1) * self.nimg] = \
                            self.cmodbimg[i, j]
            return
        self._run_device(calsource)

    def _run_device(self, calsource):
        """The integration cycle of centroidsFromPhase (cmod/src/centmodule.c:1774-1788) with only
        the readout step calling the engine (computeHll zeroes its buffer every call, :478)."""
        cc = self.centcmod
        if calsource != self._calsource:
            cc.update(centcmod_mod.CALSOURCE, calsource)
            self._calsource = calsource
        n = self.nintegrations + self.nlatency
        readout = self.curInteg == n - 1
        self.curInteg = 0 if self.curInteg + 1 >= n else self.curInteg + 1
        if not readout:
            return
        torch = self._torch
        h = cc.centstruct
        S = self.nsubx * self.nsubx
        m = S * self.nimg * self.nimg
        pois = self.addPoisson == 1 and calsource == 0
        rd = self.readnoise != 0 and calsource == 0
        k = self._readouts
        st = torch.cuda.current_stream(self.phase.device).cuda_stream
        if pois:   # the readout's noise streams, drawn on the device in stream order
            cent.drawNormalsDevice(self.seed, 2 * k, m, self._zp.data_ptr(), st)
        if rd:
            cent.drawNormalsDevice(self.seed, 2 * k + 1, m, self._zr.data_ptr(), st)
        npup = self.nsubx * self.phasesize
        phase = self.phase if self.phase.is_contiguous() else self.phase.contiguous()
        cent.runDevice(h, phase.data_ptr(), npup * npup, 1, self._dev_cents.data_ptr(), self.cmodbimg.data_ptr(),
                       self._dev_corrimg.data_ptr() if self._dev_corrimg is not None else 0,
                       self._zp.data_ptr() if pois else 0, self._zr.data_ptr() if rd else 0,
                       torch.cuda.current_stream(self.phase.device).cuda_stream)
        self._readouts += 1
        if self.imageOnly == 1:
            self.outputData.copy_(self.cmodbimg)
        elif self.imageOnly == 2:
            nb = self.nsubx * self.nimg
            self.outputData.copy_(self.cmodbimg.permute(0, 2, 1, 3).reshape(nb, nb))

    # ---- util/centroid.py:1536-1572 ----------------------------------------------------------
    def takeReference(self, control, cameraInput=None):
        """Noiseless slopes of a flat wavefront (cal_source=1), stored as reference slopes."""
        c = control["cal_source"]
        control["cal_source"] = 1
        self.centcmod.update(centcmod_mod.REFCENTS, None)
        self.phase[:] = 0
        self.runCalc(control)      # one step of the integration cycle, as the reference
        control["cal_source"] = c
        self.refCents = self._host(self.outputData).copy()
        self.centcmod.update(centcmod_mod.REFCENTS, self.refCents)
        return self.refCents

    def setRefSlopes(self, refSlopes):
        self.refCents = numpy.asarray(refSlopes).copy().astype(numpy.float32)
        self.centcmod.update(centcmod_mod.REFCENTS, self.refCents)

    # ---- util/centroid.py:1575-1633 ----------------------------------------------------------
    def takeCorrImage(self, control={"cal_source": 1}, cameraInput=None, doThreshold=0):
        """Flat-wavefront images as the correlation reference (normalised to a maximum of 1 and
        half-complex transformed for mode 1)."""
        data = None
        if self.correlationCentroiding and self.corrPattern is None:
            c = control["cal_source"]
            control["cal_source"] = 1
            self.centcmod.update(centcmod_mod.CORRELATIONCENTROIDING, 0)
            self.phase[:] = 0
            self.runCalc(control)
            self.centcmod.update(centcmod_mod.CORRELATIONCENTROIDING, self.correlationCentroiding)
            control["cal_source"] = c
            img = self._host(self.cmodbimg).copy()
            if self.corrPatternUser is None:
                self.corrPatternUser = img
            elif self.corrPatternUser.shape == img.shape:
                self.corrPatternUser[:] = img
            else:
                self.corrPatternUser[:] = 0
                s = (self.corrPatternUser.shape[-2] - img.shape[-2]) // 2
                e = s + img.shape[-2]
                self.corrPatternUser[:, :, s:e, s:e] = img
            if doThreshold:
                self.corrPatternUser = numpy.where(self.corrPatternUser < self.noiseFloor, 0,
                                                   self.corrPatternUser - self.noiseFloor)
            if self.correlationCentroiding == 1:
                self.corrPatternUser /= max(self.corrPatternUser.ravel())
                self.corrPattern = correlation.transformPSF(self.corrPatternUser)
            else:
                self.corrPattern = self.corrPatternUser.copy()
            self._engine_corr = self.corrPattern
            self.centcmod.update(centcmod_mod.CORRPATTERN, self.corrPattern)
            data = self.corrPatternUser
        return data

    def closeCmod(self):
        self.centcmod.free()
        self.centcmod = None

    def _host(self, x):
        return x.detach().cpu().numpy() if self._is_tensor(x) else numpy.asarray(x)

    @property
    def nsubaps(self):
        return self.nsubx * self.nsubx


def new_centroid(nsubx, pup, phasesize, fftsize, **kw):
    """Constructor + initMem + initialiseCmod, the sequence science/wfscent.py runs at set-up."""
    c = centroid(nsubx, pup=pup, phasesize=phasesize, fftsize=fftsize, **kw)
    c.initMem()
    c.initialiseCmod()
    return c
