"""Drop-in replacement of DASP's ``util.centcmod`` (util/centcmod.py:16-145) on the gfx950 engine.

Same constructor arguments, attributes, update codes and methods; ``cmod.cent`` is replaced by
``dasp_amd.cmod.cent``. Python 3 (the reference is Python 2).
"""
import numpy

from dasp_amd.cmod import cent

CALSOURCE = 1
SIG = 2
ADDPOISSON = 3
NCEN = 4
READNOISE = 5
READBG = 6
NOISEFLOOR = 7
SKYBRIGHTNESS = 8
PXLPOWER = 9
SPOTPSF = 10
NTHREADS = 11
OPTICALBINNING = 12
CENTWEIGHT = 13
CORRELATIONCENTROIDING = 14
CORRTHRESH = 15
CORRPATTERN = 16
CALDATA = 17
REFCENTS = 18
CALCOEFF = 19
USEBRIGHTEST = 20
INTEGSTEPS = 21   # should never be changed to something larger than the initial value
OUTPUTPHASEARR = 22


class centcmod:
    def __init__(self, nthreads, nsubx, ncen, fftsize, clipsize, nimg, phasesize, readnoise, readbg,
                 addPoisson, noiseFloor, sig, skybrightness, calsource, pxlPower, nintegrations, nlatency, seed,
                 phs, pup, spotpsf, cents, bimg, minarea, opticalBinning, centWeight, correlationCentroiding,
                 corrThresh, corrPattern, corrimg, threshType, imageOnly, useBrightest, preBinningFactor,
                 parabolicFit, gaussianFitVals, inputImage=None, subapLocation=None):
        """Wrapper for the centroid engine (util/centcmod.py:42-118). sig can be a float or a
        float32 array of nsubx*nsubx values; pup is a util.tel.Pupil-like object."""
        self.nthreads = nthreads
        self.nsubx = nsubx
        self.nsubaps = nsubx * nsubx
        self.ncen = ncen
        self.fftsize = fftsize
        self.clipsize = clipsize
        self.nimg = nimg
        self.phasesize = phasesize
        self.readnoise = readnoise
        self.readbg = readbg
        self.addPoisson = addPoisson
        self.noiseFloor = noiseFloor
        self.sig = sig
        self.skybrightness = skybrightness
        self.calsource = calsource
        self.minarea = minarea
        self.opticalBinning = opticalBinning
        if pxlPower is None:
            pxlPower = 1.
        self.pxlPower = pxlPower
        self.nintegrations = nintegrations
        self.nlatency = nlatency
        self.seed = seed
        self.phs = phs
        self.pup = pup
        self.pupfn = pup.perSubap(nsubx, vectorAlign=0)
        self.subflag = pup.getSubapFlag(nsubx, minarea).astype(numpy.int32)
        self.fracSubArea = (pup.getSubarea(nsubx) / (phasesize * phasesize)).astype(numpy.float32)
        self.spotpsf = spotpsf
        self.cents = cents
        self.bimg = bimg
        if isinstance(centWeight, str):   # will be created later
            centWeight = None
        self.centWeight = centWeight
        self.correlationCentroiding = correlationCentroiding
        self.corrThresh = corrThresh
        self.corrPattern = corrPattern
        self.corrimg = corrimg
        self.threshType = threshType
        self.imageOnly = imageOnly
        self.useBrightest = useBrightest
        self.preBinningFactor = preBinningFactor
        self.parabolicFit = parabolicFit
        if gaussianFitVals is None:
            self.gaussianFit = 0
            self.gaussianMinVal = 0.
            self.gaussianReplaceVal = 0.
        else:
            self.gaussianFit = 1
            self.gaussianMinVal = gaussianFitVals[0]
            self.gaussianReplaceVal = gaussianFitVals[1]
        self.inputImage = inputImage
        self.subapLocation = subapLocation
        if correlationCentroiding:
            if corrPattern is not None:
                if (corrPattern.dtype.char != "f" or corrPattern.shape[:2] != (nsubx, nsubx) or
                        corrPattern.shape[2] < nimg or corrPattern.shape[3] != corrPattern.shape[2]):
                    raise Exception("corrPattern wrong shape or type: %s, %s - should be (%d,%d,>=%d,>=%d), f" % (
                        str(corrPattern.shape), corrPattern.dtype.char, nsubx, nsubx, nimg, nimg))
        self.centstruct = cent.initialise(
            self.nthreads, self.nsubaps, self.ncen, self.fftsize, self.clipsize, self.nimg, self.phasesize,
            self.readnoise, self.readbg, self.addPoisson, self.noiseFloor, self.sig, self.skybrightness,
            self.calsource, self.pxlPower, self.nintegrations, self.nlatency, self.seed, self.phs, self.pupfn,
            self.spotpsf, self.cents, self.subflag, self.bimg, self.fracSubArea, self.opticalBinning,
            self.centWeight, correlationCentroiding, corrThresh, corrPattern, corrimg, threshType, imageOnly,
            useBrightest, preBinningFactor, self.parabolicFit, self.gaussianFit, self.gaussianMinVal,
            self.gaussianReplaceVal, self.inputImage, self.subapLocation)

    def run(self, calsource):
        if calsource != self.calsource:
            cent.update(self.centstruct, CALSOURCE, calsource)
            self.calsource = calsource
        return cent.run(self.centstruct)

    def runSlope(self, calsource):
        if calsource != self.calsource:
            cent.update(self.centstruct, CALSOURCE, calsource)
            self.calsource = calsource
        return cent.runSlope(self.centstruct)

    def free(self):
        if self.centstruct is not None:
            cent.free(self.centstruct)
            self.centstruct = None

    def update(self, what, val):
        """what is an update code (CALSOURCE=1 ... OUTPUTPHASEARR=22); val the new value.
        As in the reference, the Python-side attributes are not updated."""
        cent.update(self.centstruct, what, val)
