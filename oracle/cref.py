"""ctypes wrapper of the FP32 C restatement (oracle/cref/centref.c). Test infrastructure."""
import ctypes
import os

import numpy as np

from oracle.build_cref import OUT

_fp = ctypes.POINTER(ctypes.c_float)


class CrefParams(ctypes.Structure):
    _fields_ = [("nsubx", ctypes.c_int), ("n", ctypes.c_int), ("fftsize", ctypes.c_int),
                ("clipsize", ctypes.c_int), ("nimg", ctypes.c_int), ("ncen", ctypes.c_int),
                ("pupfn", ctypes.c_void_p), ("subflag", ctypes.c_void_p), ("fracSubArea", ctypes.c_void_p),
                ("sig", ctypes.c_float), ("sigArr", ctypes.c_void_p), ("skybrightness", ctypes.c_float),
                ("addPoisson", ctypes.c_int), ("calsource", ctypes.c_int), ("readnoise", ctypes.c_float),
                ("readbg", ctypes.c_float), ("noiseFloor", ctypes.c_float), ("threshType", ctypes.c_int), ("relativeFloor", ctypes.c_int),
                ("pxlPower", ctypes.c_float), ("spotpsf", ctypes.c_void_p), ("psfsize", ctypes.c_int),
                ("correlation", ctypes.c_int), ("corrPattern", ctypes.c_void_p), ("corrsize", ctypes.c_int),
                ("corrThresh", ctypes.c_float), ("parabolicFit", ctypes.c_int), ("refCents", ctypes.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(OUT):
            from oracle.build_cref import build
            build()
        _lib = ctypes.CDLL(OUT)
        _lib.cref_run.argtypes = [ctypes.POINTER(CrefParams)] + [ctypes.c_void_p] * 6 + [ctypes.c_int]
        _lib.cref_run.restype = ctypes.c_int
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


class CRef:
    """Holds the inputs (kept alive) and runs one readout of all subapertures."""

    def __init__(self, nsubx, n, fftsize, clipsize, nimg, ncen, pupfn, subflag, frac, sig=1.0, sigArr=None,
                 skybrightness=0., addPoisson=0, calsource=0, readnoise=0., readbg=0., noiseFloor=0.,
                 threshType=0, pxlPower=1., spotpsf=None, corrPattern=None, corrThresh=0., parabolicFit=0,
                 refCents=None, relativeFloor=False):
        f32 = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
        self.keep = dict(pupfn=f32(pupfn), subflag=np.ascontiguousarray(subflag, np.int32), frac=f32(frac),
                         sigArr=f32(sigArr), spot=f32(spotpsf), corr=f32(corrPattern), ref=f32(refCents))
        k = self.keep
        self.S = nsubx * nsubx
        self.nimg = nimg
        self.C = corrPattern.shape[-1] if corrPattern is not None else nimg
        self.p = CrefParams(nsubx, n, fftsize, clipsize, nimg, ncen, _p(k["pupfn"]), _p(k["subflag"]),
                            _p(k["frac"]), sig, _p(k["sigArr"]), skybrightness, addPoisson, calsource, readnoise,
                            readbg, noiseFloor, threshType, int(bool(relativeFloor)), pxlPower, _p(k["spot"]),
                            spotpsf.shape[-1] if spotpsf is not None else 0, 1 if corrPattern is not None else 0,
                            _p(k["corr"]), self.C, corrThresh, parabolicFit, _p(k["ref"]))

    def run(self, phs, zpois=None, zread=None, nthreads=1):
        phs = np.ascontiguousarray(phs, np.float32)
        zp = None if zpois is None else np.ascontiguousarray(zpois, np.float32)
        zr = None if zread is None else np.ascontiguousarray(zread, np.float32)
        cents = np.zeros((self.S, 2), np.float32)
        bimg = np.zeros((self.S, self.nimg, self.nimg), np.float32)
        corr = np.zeros((self.S, self.C, self.C), np.float32) if self.keep["corr"] is not None else None
        rc = lib().cref_run(ctypes.byref(self.p), _p(phs), _p(zp), _p(zr), _p(cents), _p(bimg), _p(corr),
                            int(nthreads))
        if rc != 0:
            raise RuntimeError("cref_run failed")
        return cents, bimg, corr
