"""Shared builders for the GPU parity tests: one description -> (engine via centcmod, oracle)."""
import sys

import numpy as np

from dasp_amd.util import centcmod
from dasp_amd.util.tel import obstructed_pupil
from oracle import centoracle as O


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class Case:
    """A centcmod instance and the matching oracle Params for the same inputs."""

    def __init__(self, nsubx, n, nfft, nimg, obs=0.28, clip=None, ncen=None, sig=1.0, readnoise=0.,
                 readbg=0., addPoisson=0, noiseFloor=0., skybrightness=0., calsource=0, pxlPower=1.,
                 nintegrations=1, nlatency=0, seed=1, spotpsf=None, opticalBinning=0, centWeight=None,
                 correlationCentroiding=0, corrThresh=0., corrPattern=None, threshType=0, imageOnly=0,
                 useBrightest=0, preBinningFactor=1, parabolicFit=0, gaussianFitVals=None,
                 inputImage=None, subapLocation=None, phs=None, pupil=None, form=None, relativeFloor=False):
        self.nsubx, self.n, self.nfft, self.nimg = nsubx, n, nfft, nimg
        self.clip = clip or nfft
        self.ncen = ncen or nimg
        self.pupil = pupil or obstructed_pupil(nsubx, n, obs, 0.5)
        npup = nsubx * n
        self.phs = np.zeros((npup, npup), np.float32) if phs is None else phs
        self.cents = np.zeros((nsubx, nsubx, 2), np.float32)
        self.bimg = np.zeros((nsubx, nsubx, nimg, nimg), np.float32)
        self.corrimg = None
        hc = None
        if corrPattern is not None:
            hc = O.transform_psf(corrPattern) if correlationCentroiding == 1 else corrPattern.astype(np.float32)
            self.corrimg = np.zeros(hc.shape, np.float32)
        self.hc = hc
        self.cc = centcmod.centcmod(
            1, nsubx, self.ncen, nfft, self.clip, nimg, n, readnoise, readbg, addPoisson, noiseFloor, sig,
            skybrightness, calsource, pxlPower, nintegrations, nlatency, seed, self.phs, self.pupil, spotpsf,
            self.cents, self.bimg, 0.5, opticalBinning, centWeight, correlationCentroiding, corrThresh, hc,
            self.corrimg, threshType, imageOnly, useBrightest, preBinningFactor, parabolicFit, gaussianFitVals,
            inputImage, subapLocation)
        if relativeFloor:   # engine option: threshType 0 with the floor relative to the image maximum
            from dasp_amd.cmod import cent
            cent.setRelativeFloor(self.cc.centstruct, True)
        if form is not None:
            from dasp_amd.cmod import cent
            cent.setForm(self.cc.centstruct, form % 10)
            cent.setFastTrig(self.cc.centstruct, form >= 10)   # 0..3: accurate sincosf, 10..13: fast
        gf = 0 if gaussianFitVals is None else 1
        sigarr = sig if isinstance(sig, np.ndarray) else None
        skyarr = skybrightness if isinstance(skybrightness, np.ndarray) else None
        ubarr = useBrightest if isinstance(useBrightest, np.ndarray) else None
        self.params = O.Params(
            nsubx, n, nfft, self.clip, nimg, self.ncen, self.cc.pupfn, self.cc.subflag, self.cc.fracSubArea,
            readnoise=readnoise, readbg=readbg, addPoisson=addPoisson, noiseFloor=noiseFloor,
            sig=0.0 if sigarr is not None else float(sig), sigArr=sigarr,
            skybrightness=0.0 if skyarr is not None else float(skybrightness), skybrightnessArr=skyarr,
            calsource=calsource, pxlPower=pxlPower, spotpsf=spotpsf, opticalBinning=opticalBinning,
            centWeight=centWeight, correlationCentroiding=correlationCentroiding, corrThresh=corrThresh,
            corrPattern=hc, threshType=threshType, imageOnly=imageOnly,
            useBrightest=0 if ubarr is not None else int(useBrightest), useBrightestArr=ubarr,
            preBinningFactor=preBinningFactor, parabolicFit=parabolicFit, gaussianFit=gf,
            gaussianMinVal=0. if gf == 0 else gaussianFitVals[0],
            gaussianReplaceVal=0. if gf == 0 else gaussianFitVals[1], relativeFloor=relativeFloor)
        self.seed = seed
        self.readouts = 0

    def noise(self):
        """The normals the engine draws for its next readout (streams 2k and 2k+1)."""
        from dasp_amd.cmod import cent
        S = self.nsubx * self.nsubx
        m2 = self.nimg * self.nimg
        zp = cent.drawNormals(self.seed, 2 * self.readouts, S * m2).reshape(S, self.nimg, self.nimg)
        zr = cent.drawNormals(self.seed, 2 * self.readouts + 1, S * m2).reshape(S, self.nimg, self.nimg)
        return zp, zr

    def run_both(self):
        zp, zr = self.noise()
        bimg_before = self.bimg.copy()
        self.cc.run(self.params.calsource)
        self.readouts += 1
        oc, ob, ocorr = O.run_phase(self.params, self.phs, zp, zr, bimg_in=bimg_before)
        return oc, ob, ocorr

    def free(self):
        self.cc.free()


def report(name, got, want, active):
    d = np.abs(got.reshape(-1, 2) - want.reshape(-1, 2))[active]
    say("%s: max|d|=%.3g px, rms=%.3g px over %d subaps" % (name, d.max(), np.sqrt((d ** 2).mean()), active.sum()))
    return d
