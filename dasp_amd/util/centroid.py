"""The hot-path part of util/centroid.py's ``centroid`` class over the gfx950 engine.

Covers what sits on the phase -> slopes path: the derived sizes and integration counts of the
constructor (util/centroid.py:115-340), ``initMem`` (:442-493, output and image buffers),
``initialiseCmod`` (:495-520), the ``runCalc`` / ``runCmod`` / ``runSlopeCalc`` dispatch
(:547-615) with its reference-slope (``doref``) rule, ``takeReference`` / ``setRefSlopes``
(:1536-1572), ``takeCorrImage`` (:1575-1633), ``takeCentWeight`` (:1661-1679) and the
linearisation set-up ``calibrateSHS`` / ``calibrateSHSUnique`` / ``calibrateSHSIdentical`` /
``makeCalibrationCoeffs`` / ``applyCalibration*`` (:1162-1533). The pure-Python imaging route
(``runPy``) is not part of it.

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
                 linearSteps=None, stepRangeFrac=1., calNCoeff=0, printLinearisationForcing=0,
                 inputImage=None, subapLocation=None, device=None):
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
        self.linearSteps = linearSteps
        self.stepRangeFrac = stepRangeFrac
        self.calNCoeff = calNCoeff
        self.printLinearisationForcing = printLinearisationForcing
        self.magicCentroiding = 0
        self.inputImage = inputImage
        self.subapLocation = subapLocation
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
            self.preBinningFactor, self.parabolicFit, self.gaussianFitVals, self.inputImage, self.subapLocation)
        self._calsource = calsource

    # ---- util/centroid.py:547-615 ------------------------------------------------------------
    def _engine_refs(self):
        """True when the engine subtracts the reference slopes (no calibration, or calibration done
        in the engine); otherwise the Python-side identical-subaperture calibration runs after the
        engine and the reference slopes are subtracted after it (util/centroid.py:568, :1545)."""
        return (self.linearSteps is None or self.psf is not None or self.correlationCentroiding != 0
                or self.calNCoeff != 0)

    def _finish(self):
        """Calibration and reference slopes after an engine run (util/centroid.py:574-578)."""
        if self.linearSteps is not None:
            self.applyCalibration()
        if not self._engine_refs() and self.refCents is not None:
            if self._is_tensor(self.outputData):
                self.outputData -= self._torch.as_tensor(self.refCents, device=self.outputData.device)
            else:
                self.outputData -= self.refCents

    def runSlopeCalc(self, control):
        """Image input (util/centroid.py:547-558): the detector image in inputImage -> slopes."""
        if self.inputImage is None:
            raise ValueError("runSlopeCalc needs a centroid object constructed with inputImage")
        if self.device is not None:
            raise ValueError("image input runs on the host route (device=None)")
        self.centcmod.runSlope(control["cal_source"])
        if self.imageOnly != 0:
            raise Exception("no point getting here!  If you want the image only, don't use this module.")
        self._finish()

    def runCalc(self, control={"cal_source": 0}):
        if self.phaseMultiplier != 1:
            self.phase *= self.phaseMultiplier       # in place, as the reference
        self.runCmod(control["cal_source"])
        self._finish()

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
        """Noiseless slopes of a flat wavefront (cal_source=1), or of the current camera image,
        stored as reference slopes; handed to the engine only when it also does the calibration
        (or there is none)."""
        c = control["cal_source"]
        control["cal_source"] = 1
        if self._engine_refs():
            self.centcmod.update(centcmod_mod.REFCENTS, None)
        if cameraInput is not None or self.inputImage is not None:
            if cameraInput is not None:
                self.inputImage[:] = cameraInput
            self.runSlopeCalc(control)
        else:
            self.phase[:] = 0
            self.runCalc(control)      # one step of the integration cycle, as the reference
        control["cal_source"] = c
        self.refCents = self._host(self.outputData).copy()
        if self._engine_refs():
            self.centcmod.update(centcmod_mod.REFCENTS, self.refCents)
        return self.refCents

    def setRefSlopes(self, refSlopes):
        self.refCents = numpy.asarray(refSlopes).copy().astype(numpy.float32)
        if self._engine_refs():
            self.centcmod.update(centcmod_mod.REFCENTS, self.refCents)

    # ---- util/centroid.py:1661-1679 ----------------------------------------------------------
    def takeCentWeight(self, control):
        """A string centWeight ("make") becomes the noiseless flat-wavefront image of every
        sub-aperture, used as the CoG weighting (update CENTWEIGHT)."""
        if isinstance(self.centWeight, str) and self.corrPattern is None:
            c = control["cal_source"]
            control["cal_source"] = 1
            steps, self.linearSteps = self.linearSteps, None
            self.phase[:] = 0
            self.runCalc(control)
            control["cal_source"] = c
            self.linearSteps = steps
            self.centWeight = self._host(self.cmodbimg).copy()
            self.centcmod.update(centcmod_mod.CENTWEIGHT, self.centWeight)

    # ---- linearisation set-up, util/centroid.py:1162-1533 ---------------------------------------
    def _step_list(self, steps):
        return ((numpy.arange(steps) - steps / 2. + 0.5) / (steps / 2. - 0.5) * numpy.pi *
                self.stepRangeFrac).astype(numpy.float32)

    def _sweep(self, control, steps):
        """Centroids of a sweep of phase tilts, noise off (Poisson and read noise updated to 0
        and restored after), cal_source 0 so that backgrounds and thresholds stay as in use."""
        data = numpy.zeros((2, self.nsubx, self.nsubx, steps), numpy.float32)
        stepList = self._step_list(steps)
        c = control["cal_source"]
        control["cal_source"] = 0
        self.centcmod.update(centcmod_mod.ADDPOISSON, 0)
        self.centcmod.update(centcmod_mod.READNOISE, 0.)
        npup = self.nsubx * self.phasesize
        for i in range(steps):
            tilt = (-numpy.arange(npup) * stepList[i]).astype(numpy.float32)
            self._set_phase(numpy.broadcast_to(tilt[None, :], (npup, npup)))
            self.runCalc(control)
            data[0, :, :, i] = self._host(self.outputData[:, :, 0])
            self._set_phase(numpy.broadcast_to(tilt[:, None], (npup, npup)))
            self.runCalc(control)
            data[1, :, :, i] = self._host(self.outputData[:, :, 1])
        self.centcmod.update(centcmod_mod.ADDPOISSON, int(self.centcmod.addPoisson))
        self.centcmod.update(centcmod_mod.READNOISE, float(self.centcmod.readnoise))
        control["cal_source"] = c
        return data, stepList

    def _set_phase(self, a):
        if self._is_tensor(self.phase):
            self.phase.copy_(self._torch.from_numpy(numpy.ascontiguousarray(a, numpy.float32)))
        else:
            self.phase[:] = a

    def calibrateSHSUnique(self, control={"cal_source": 1}):
        """Every sub-aperture calibrated on its own (util/centroid.py:1162-1247): a sweep of
        linearSteps tilts, the single-valued bounds argmin..argmax per sub-aperture, and points
        that break the monotonic order replaced by the mean of their neighbours (nearest-neighbour
        calibration only, calNCoeff == 0)."""
        if self.linearSteps is None:
            return
        steps = self.linearSteps
        self.linearSteps = None
        self.calibrateData, self.calibrateSteps = self._sweep(control, steps)
        self.linearSteps = steps
        cd = self.calibrateData
        self.calibrateBounds = numpy.zeros((2, self.nsubx, self.nsubx, 2), numpy.int32)
        self.calibrateBounds[..., 0] = numpy.argmin(cd, axis=3)
        self.calibrateBounds[..., 1] = numpy.argmax(cd, axis=3)
        forced, maxShift = 0, 0.
        if self.calNCoeff == 0:
            for d in range(2):
                for i in range(self.nsubx):
                    for j in range(self.nsubx):
                        for k in range(self.calibrateBounds[d, i, j, 0], self.calibrateBounds[d, i, j, 1]):
                            if cd[d, i, j, k] > cd[d, i, j, k + 1]:
                                val = (cd[d, i, j, k - 1] + cd[d, i, j, k + 1]) / 2.
                                if self.printLinearisationForcing:
                                    print("INFORMATION:**centroid**: Forcing SHS calibration for point (%d,%d) "
                                          "step %d from %g to %g" % (i, j, k, cd[d, i, j, k], val))
                                forced += 1
                                maxShift = max(maxShift, abs(cd[0, i, j, k] - val))   # (the reference's index)
                                cd[d, i, j, k] = val
        self.linearPointsForced = forced

    def applyCalibrationUnique(self, data=None):
        """util/centroid.py:1249-1296, the Python route of the engine's own calibration."""
        data = self.outputData if data is None else data
        host = self._host(data).copy()
        for i in range(self.nsubx):
            for j in range(self.nsubx):
                if not self.subflag[i, j]:
                    continue
                if self.calNCoeff == 0:
                    for d in range(2):
                        f, t = self.calibrateBounds[d, i, j]
                        host[i, j, d] = numpy.interp([host[i, j, d]], self.calibrateData[d, i, j, f:t + 1],
                                                     self.calibrateSteps[f:t + 1])[0]
                else:
                    for d in range(2):
                        c = float(host[i, j, d])
                        host[i, j, d] = sum(float(self.calCoeff[i, j, d, k]) * c ** k for k in range(self.calNCoeff))
        self._put(data, host)

    def calibrateSHSIdentical(self, control={"cal_source": 1}):
        """Sub-apertures with identical pupil functions share one calibration (util/centroid.py
        :1298-1410); applied in Python after the engine (applyCalibrationIdentical)."""
        if self.linearSteps is None:
            return
        steps = self.linearSteps
        self.linearSteps = None
        n = self.phasesize
        pupfn = self.pup.fn if hasattr(self.pup, "fn") else self.pup
        pupfn = numpy.asarray(pupfn)
        masks = {0: []}
        self.calMaskType = numpy.zeros((self.nsubx, self.nsubx), numpy.int32)
        for i in range(self.nsubx):
            for j in range(self.nsubx):
                if self.subflag[i, j]:
                    win = pupfn[i * n:(i + 1) * n, j * n:(j + 1) * n]
                    typ = None
                    for key in masks:   # key 0 is the reference's empty placeholder type
                        if key != 0 and numpy.array_equal(masks[key], win):
                            typ = key
                            break
                    if typ is None:
                        typ = max(masks) + 1
                        masks[typ] = win
                    self.calMaskType[i, j] = typ
        self.calibrateData, self.calibrateSteps = self._sweep(control, steps)
        self.linearSteps = steps
        self.calibrateBounds = numpy.zeros((2, self.nsubx, self.nsubx, 2), numpy.int32)
        self.calDataDict = {}
        forced, maxShift = 0, 0.
        for i in range(self.nsubx):
            for j in range(self.nsubx):
                if not self.subflag[i, j]:
                    continue
                typ = self.calMaskType[i, j]
                if typ not in self.calDataDict:
                    lists = []
                    for d in range(2):
                        f = numpy.argmin(self.calibrateData[d, i, j])
                        t = numpy.argmax(self.calibrateData[d, i, j]) + 1
                        lists += [self.calibrateData[d, i, j, f:t].copy(), self.calibrateSteps[f:t]]
                    cd = CalData(*lists)
                    for attr_c, attr_r in (("xc", "xr"), ("yc", "yr")):
                        vc = getattr(cd, attr_c)
                        for k in range(vc.shape[0] - 1):
                            if vc[k] >= vc[k + 1]:
                                val = vc[k + 1] - (vc[k + 2] - vc[k + 1]) if k == 0 else (vc[k - 1] + vc[k + 1]) / 2.
                                forced += 1
                                maxShift = max(maxShift, abs(val - vc[k]))
                                vc[k] = val
                        notsame = numpy.nonzero(vc[1:] != vc[:-1])[0]
                        setattr(cd, attr_c, vc[notsame])
                        setattr(cd, attr_r, getattr(cd, attr_r)[notsame])
                    self.calDataDict[typ] = cd
                self.calDataDict[typ].indx.append(i * self.nsubx + j)
        for cd in self.calDataDict.values():
            cd.xindx = numpy.array(cd.indx).astype(numpy.int32) * 2
            cd.yindx = cd.xindx + 1
        self.linearPointsForced = forced

    def applyCalibrationIdentical(self, data=None):
        """util/centroid.py:1412-1475: per pupil type, numpy.interp of the slopes on its table."""
        data = self.outputData if data is None else data
        rdata = self._host(data).astype(numpy.float32).ravel().copy()
        for cd in self.calDataDict.values():
            x = numpy.take(rdata, cd.xindx)
            y = numpy.take(rdata, cd.yindx)
            if cd.xc.size != 0:
                rdata[cd.xindx] = numpy.interp(x, cd.xc, cd.xr).astype(numpy.float32)
            if cd.yc.size != 0:
                rdata[cd.yindx] = numpy.interp(y, cd.yc, cd.yr).astype(numpy.float32)
        self._put(data, rdata.reshape(self._host(data).shape))

    def makeCalibrationCoeffs(self):
        """Polynomial calibration per sub-aperture (util/centroid.py:1477-1489)."""
        self.calCoeff = numpy.zeros((self.nsubx, self.nsubx, 2, self.calNCoeff), numpy.float32)
        for i in range(self.nsubx):
            for j in range(self.nsubx):
                if self.subflag[i, j]:
                    f, t = self.calibrateBounds[0, i, j]
                    for d in range(2):
                        self.calCoeff[i, j, d] = numpy.polyfit(self.calibrateData[d, i, j, f:t], self.calibrateSteps[f:t],
                                                               self.calNCoeff - 1)[::-1]

    def calibrateSHS(self, control={"cal_source": 1}):
        """util/centroid.py:1492-1504: identical-pupil calibration in Python when the engine does no
        calibration of its own, else the per-sub-aperture tables (CALDATA) or polynomials (CALCOEFF)
        handed to the engine."""
        if self.linearSteps is None:
            return
        if self.psf is None and self.correlationCentroiding == 0 and self.calNCoeff == 0:
            self.calibrateSHSIdentical(control)
        else:
            self.calibrateSHSUnique(control)
            if self.calNCoeff != 0:
                self.makeCalibrationCoeffs()
                self.centcmod.update(centcmod_mod.CALCOEFF, self.calCoeff)
            else:
                self.centcmod.update(centcmod_mod.CALDATA, (self.calibrateData, self.calibrateBounds,
                                                            self.calibrateSteps))

    def applyCalibration(self, data=None):
        """util/centroid.py:1506-1511: only the identical-pupil calibration runs in Python; the others
        are done by the engine."""
        if self.psf is None and self.correlationCentroiding == 0 and self.calNCoeff == 0:
            self.applyCalibrationIdentical(data)

    def _put(self, data, host):
        if self._is_tensor(data):
            data.copy_(self._torch.from_numpy(numpy.ascontiguousarray(host, numpy.float32)))
        else:
            data[:] = host

    @property
    def subflag(self):
        return self.centcmod.subflag.reshape(self.nsubx, self.nsubx)

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


class CalData:
    """Calibration table of one pupil type (util/centroid.py:1935-1943)."""

    def __init__(self, xc, xr, yc, yr):
        self.xc = xc
        self.xr = xr
        self.yc = yc
        self.yr = yr
        self.indx = []
        self.xindx = None
        self.yindx = None


def new_centroid(nsubx, pup, phasesize, fftsize, **kw):
    """Constructor + initMem + initialiseCmod, the sequence science/wfscent.py runs at set-up."""
    c = centroid(nsubx, pup=pup, phasesize=phasesize, fftsize=fftsize, **kw)
    c.initMem()
    c.initialiseCmod()
    return c
