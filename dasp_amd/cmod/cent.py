"""``cent`` - the gfx950 replacement of DASP's CPython module ``cmod.cent``.

Same functions, argument order and behaviour as cmod/src/centmodule.c:3313-3321:
``initialise(41 args) -> int handle``, ``run(handle) -> float seconds``,
``runSlope(handle) -> float``, ``update(handle, code, value)``, ``free(handle)`` and
``testcorrelation(img, corr)``. The work happens in ``libdasp_cent.so`` (HIP kernels for gfx950)
through its C ABI (include/dasp_cent.h); this module only converts Python arguments.

As in the reference, the numpy arrays passed in are borrowed: the caller keeps them alive, the
phase is re-read at every ``run`` and ``cents``/``bimg``/``corrimg`` are written in place.
Failures raise ``cent.error`` (the reference printed and returned NULL).
"""
import ctypes
import os

import numpy

# DASP_CENT_LIB: path of another build of the library (default: the in-tree build)
_LIBPATH = os.environ.get("DASP_CENT_LIB") or os.path.join(
    os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "libdasp_cent.so")


class error(Exception):
    """Raised when the engine rejects a call (message from dasp_cent_last_error)."""


class _Array(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("ndim", ctypes.c_int), ("dims", ctypes.c_long * 4)]


class _Params(ctypes.Structure):
    _fields_ = [
        ("nthreads", ctypes.c_int), ("nsubaps", ctypes.c_int), ("ncen", ctypes.c_int),
        ("fftsize", ctypes.c_int), ("clipsize", ctypes.c_int), ("nimg", ctypes.c_int),
        ("phasesize", ctypes.c_int), ("readnoise", ctypes.c_float), ("readbg", ctypes.c_float),
        ("addPoisson", ctypes.c_int), ("noiseFloor", ctypes.c_float), ("sig", ctypes.c_float),
        ("sigArr", _Array), ("skybrightness", ctypes.c_float), ("skybrightnessArr", _Array),
        ("calsource", ctypes.c_int), ("pxlPower", ctypes.c_float), ("nintegrations", ctypes.c_int),
        ("nlatency", ctypes.c_int), ("seed", ctypes.c_ulong), ("phs", _Array), ("pupfn", _Array),
        ("spotpsf", _Array), ("cents", _Array), ("subflag", _Array), ("bimg", _Array),
        ("fracSubArea", _Array), ("opticalBinning", ctypes.c_int), ("centWeight", _Array),
        ("correlationCentroiding", ctypes.c_int), ("corrThresh", ctypes.c_float),
        ("corrPattern", _Array), ("corrimg", _Array), ("threshType", ctypes.c_int),
        ("imageOnly", ctypes.c_int), ("useBrightest", ctypes.c_int), ("useBrightestArr", _Array),
        ("preBinningFactor", ctypes.c_int), ("parabolicFit", ctypes.c_int),
        ("gaussianFit", ctypes.c_int), ("gaussianMinVal", ctypes.c_float),
        ("gaussianReplaceVal", ctypes.c_float), ("inputImage", _Array), ("subapLocation", _Array),
    ]


class _BatchItem(ctypes.Structure):
    _fields_ = [("handle", ctypes.c_void_p), ("phs", ctypes.c_void_p), ("phs_stride", ctypes.c_long),
                ("nframes", ctypes.c_int), ("cents", ctypes.c_void_p), ("bimg", ctypes.c_void_p),
                ("corrimg", ctypes.c_void_p), ("zpois", ctypes.c_void_p), ("zread", ctypes.c_void_p)]


_lib = None


def library():
    """Load libdasp_cent.so (built by dasp_amd.build); raise if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            raise ImportError("libdasp_cent.so not built (run python -m dasp_amd.build): %s" % _LIBPATH)
        # One HIP runtime per process: PyTorch ships its own libamdhip64 (same soname). Loading it
        # first makes this library bind to it, so torch tensors and the engine share the device
        # context; loading the system runtime first would leave torch unable to initialise HIP.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(_LIBPATH)
        vp, ip, fp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_float)
        lib.dasp_cent_last_error.restype = ctypes.c_char_p
        lib.dasp_cent_abi_version.restype = ctypes.c_int
        lib.dasp_cent_initialise.argtypes = [ctypes.POINTER(_Params), ctypes.POINTER(vp)]
        lib.dasp_cent_run.argtypes = [vp, fp]
        lib.dasp_cent_run_slope.argtypes = [vp, fp]
        lib.dasp_cent_update_int.argtypes = [vp, ip, ctypes.c_long]
        lib.dasp_cent_update_float.argtypes = [vp, ip, ctypes.c_double]
        lib.dasp_cent_update_array.argtypes = [vp, ip, ctypes.POINTER(_Array)]
        lib.dasp_cent_update_caldata.argtypes = [vp] + [ctypes.POINTER(_Array)] * 3
        lib.dasp_cent_free.argtypes = [vp]
        lib.dasp_cent_testcorrelation.argtypes = [vp, vp, ip]
        lib.dasp_cent_set_form.argtypes = [vp, ip]
        lib.dasp_cent_run_device.argtypes = [vp, vp, ctypes.c_long, ip, vp, vp, vp, vp, vp, vp]
        lib.dasp_cent_draw_normals.argtypes = [ctypes.c_ulong, ctypes.c_ulong, ctypes.c_long, vp]
        lib.dasp_cent_draw_normals_device.argtypes = [ctypes.c_ulong, ctypes.c_ulong, ctypes.c_long, vp, vp]
        lib.dasp_cent_hc_to_spectrum.argtypes = [vp, ip, vp]
        lib.dasp_cent_info.argtypes = [vp, ctypes.POINTER(ip), ctypes.POINTER(ip), ctypes.POINTER(ip)]
        lib.dasp_cent_set_timing.argtypes = [vp, ip]
        lib.dasp_cent_set_fast_trig.argtypes = [vp, ip]
        lib.dasp_cent_kernel_time.argtypes = [vp, ip, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long)]
        lib.dasp_cent_batch_create.argtypes = [ctypes.POINTER(_BatchItem), ip, ctypes.POINTER(vp)]
        lib.dasp_cent_batch_tune.argtypes = [vp, vp, ip]
        lib.dasp_cent_batch_run.argtypes = [vp, vp]
        lib.dasp_cent_batch_launches.argtypes = [vp]
        lib.dasp_cent_batch_launch_info.argtypes = [vp, ip, ctypes.POINTER(ip), ctypes.POINTER(ip),
                                                    ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ip),
                                                    ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ip),
                                                    ctypes.POINTER(ctypes.c_float)]
        lib.dasp_cent_batch_others.argtypes = [vp]
        lib.dasp_cent_batch_set_config.argtypes = [vp, ip, ip, ctypes.c_long, ip]
        lib.dasp_cent_batch_free.argtypes = [vp]
        for name in ("dasp_cent_batch_create", "dasp_cent_batch_tune", "dasp_cent_batch_run",
                     "dasp_cent_batch_launches", "dasp_cent_batch_launch_info", "dasp_cent_batch_others",
                     "dasp_cent_batch_free", "dasp_cent_batch_set_config"):
            getattr(lib, name).restype = ctypes.c_int
        for name in ("dasp_cent_initialise", "dasp_cent_run", "dasp_cent_run_slope", "dasp_cent_update_int",
                     "dasp_cent_update_float", "dasp_cent_update_array", "dasp_cent_update_caldata",
                     "dasp_cent_free", "dasp_cent_testcorrelation", "dasp_cent_set_form",
                     "dasp_cent_run_device", "dasp_cent_draw_normals", "dasp_cent_draw_normals_device",
                     "dasp_cent_hc_to_spectrum",
                     "dasp_cent_info", "dasp_cent_set_timing", "dasp_cent_kernel_time",
                     "dasp_cent_set_fast_trig"):
            getattr(lib, name).restype = ctypes.c_int
        _lib = lib
    return _lib


def _check(rc):
    if rc != 0:
        raise error(library().dasp_cent_last_error().decode(errors="replace") or ("error %d" % rc))


def _is_f32(a):
    return isinstance(a, numpy.ndarray) and a.dtype == numpy.float32 and a.flags.c_contiguous


def _is_i32(a):
    return (isinstance(a, numpy.ndarray) and a.dtype.kind == "i" and a.dtype.itemsize == 4
            and a.flags.c_contiguous)


def _arr(a, what, kind="f", allow_none=True):
    """A dasp_cent_array over a numpy buffer (checkFloatContigArr/checkIntContigArr, :2132-2149)."""
    r = _Array()
    if a is None:
        if not allow_none:
            raise error("centmodule - %s must not be None" % what)
        return r
    if not isinstance(a, numpy.ndarray):
        if allow_none:
            return r   # the reference treats a non-array object as "not given" (PyArray_Check)
        raise error("centmodule - %s must be an array" % what)
    if (kind == "f" and not _is_f32(a)) or (kind == "i" and not _is_i32(a)):
        raise error("centmodule - %s must be %s and contiguous" % (what, "float32" if kind == "f" else "int32"))
    if a.ndim > 4:
        raise error("centmodule - %s has too many dimensions" % what)
    r.data = a.ctypes.data
    r.ndim = a.ndim
    for i, d in enumerate(a.shape):
        r.dims[i] = d
    return r


class _Handle:
    __slots__ = ("ptr", "keep")


_handles = {}


def _ptr(handle):
    try:
        return _handles[handle].ptr
    except KeyError:
        raise error("invalid centstruct handle %r" % (handle,))


def initialise(nthreads, nsubaps, ncen, fftsize, clipsize, nimg, phasesize, readnoise, readbg,
               addPoisson, noiseFloor, sig, skybrightness, calsource, pxlPower, nintegrations,
               nlatency, seed, phs, pupfn, spotpsf, cents, subflag, bimg, fracSubArea,
               opticalBinning, centWeight, correlationCentroiding, corrThresh, corrPattern, corrimg,
               threshType, imageOnly, useBrightest, preBinningFactor, parabolicFit, gaussianFit,
               gaussianMinVal, gaussianReplaceVal, inputImage, subapLocation):
    """cent.initialise (cmod/src/centmodule.c:2858-3161); returns an integer handle."""
    p = _Params()
    p.nthreads, p.nsubaps, p.ncen = int(nthreads), int(nsubaps), int(ncen)
    p.fftsize, p.clipsize, p.nimg, p.phasesize = int(fftsize), int(clipsize), int(nimg), int(phasesize)
    p.readnoise, p.readbg = float(readnoise), float(readbg)
    p.addPoisson, p.noiseFloor = int(addPoisson), float(noiseFloor)
    if isinstance(sig, numpy.ndarray):
        p.sigArr = _arr(sig, "sig array")
    elif isinstance(sig, (float, numpy.floating)):
        p.sig = float(sig)
    else:
        raise error("Error: centmodule - sig object must be float or float array")
    if isinstance(skybrightness, numpy.ndarray):
        p.skybrightnessArr = _arr(skybrightness, "skybrightness")
    else:
        p.skybrightness = float(skybrightness)
    p.calsource, p.pxlPower = int(calsource), float(pxlPower)
    p.nintegrations, p.nlatency, p.seed = int(nintegrations), int(nlatency), int(seed)
    p.phs = _arr(phs, "phs")
    p.pupfn = _arr(pupfn, "pupfn", allow_none=False)
    p.spotpsf = _arr(spotpsf, "spotpsf")
    p.cents = _arr(cents, "cents", allow_none=False)
    p.subflag = _arr(subflag, "subflag", "i", allow_none=False)
    p.bimg = _arr(bimg, "bimg", allow_none=False)
    p.fracSubArea = _arr(fracSubArea, "fracSubArea", allow_none=False)
    p.opticalBinning = int(opticalBinning)
    p.centWeight = _arr(centWeight, "centWeight")
    p.correlationCentroiding, p.corrThresh = int(correlationCentroiding), float(corrThresh)
    p.corrPattern = _arr(corrPattern, "corrPattern")
    p.corrimg = _arr(corrimg, "corrimg")
    p.threshType, p.imageOnly = int(threshType), int(imageOnly)
    if isinstance(useBrightest, numpy.ndarray):
        p.useBrightestArr = _arr(useBrightest, "useBrightest", "i")
    elif isinstance(useBrightest, (int, numpy.integer)):
        p.useBrightest = int(useBrightest)
    else:
        raise error("Error: centmodule - useBrightest must be int or int array")
    p.preBinningFactor, p.parabolicFit, p.gaussianFit = int(preBinningFactor), int(parabolicFit), int(gaussianFit)
    p.gaussianMinVal, p.gaussianReplaceVal = float(gaussianMinVal), float(gaussianReplaceVal)
    p.inputImage = _arr(inputImage, "inputImage")
    p.subapLocation = _arr(subapLocation, "subapLocation", "i")
    out = ctypes.c_void_p()
    _check(library().dasp_cent_initialise(ctypes.byref(p), ctypes.byref(out)))
    h = _Handle()
    h.ptr = out
    h.keep = [phs, pupfn, spotpsf, cents, subflag, bimg, fracSubArea, centWeight, corrPattern,
              corrimg, inputImage, subapLocation, sig, skybrightness, useBrightest]
    key = int(out.value)
    _handles[key] = h
    return key


def run(handle):
    """cent.run (cmod/src/centmodule.c:3188-3229): returns the elapsed time in seconds."""
    t = ctypes.c_float()
    _check(library().dasp_cent_run(_ptr(handle), ctypes.byref(t)))
    return t.value


def runSlope(handle):
    """cent.runSlope (cmod/src/centmodule.c:3230-3272)."""
    t = ctypes.c_float()
    _check(library().dasp_cent_run_slope(_ptr(handle), ctypes.byref(t)))
    return t.value


_ARRAY_CODES = {8: "skybrightness", 10: "spotpsf", 13: "centWeight", 16: "corrPattern", 18: "refCents",
                19: "calCoeff", 22: "outputPhaseArr"}


def update(handle, code, obj):
    """cent.update (cmod/src/centmodule.c:2573-2856): dispatch on the Python type of obj."""
    lib = library()
    ptr = _ptr(handle)
    code = int(code)
    if code == 17:   # CALDATA: None or (calData, calBounds, calSteps)
        if obj is None:
            _check(lib.dasp_cent_update_caldata(ptr, None, None, None))
            return None
        if not isinstance(obj, tuple) or len(obj) < 3:
            raise error("Error: centmodule - calibration stuff not a tuple")
        a = [_arr(obj[0], "calData", allow_none=False), _arr(obj[1], "calBounds", "i", allow_none=False),
             _arr(obj[2], "calSteps", allow_none=False)]
        _check(lib.dasp_cent_update_caldata(ptr, *[ctypes.byref(x) for x in a]))
        _handles[handle].keep.append(obj)
        return None
    if isinstance(obj, numpy.ndarray):
        kind = "i" if code == 20 else "f"
        a = _arr(obj, _ARRAY_CODES.get(code, "value"), kind, allow_none=False)
        _check(lib.dasp_cent_update_array(ptr, code, ctypes.byref(a)))
        _handles[handle].keep.append(obj)
        return None
    if obj is None or (code in _ARRAY_CODES and not isinstance(obj, (int, float, numpy.number))):
        a = _Array()
        _check(lib.dasp_cent_update_array(ptr, code, ctypes.byref(a)))
        return None
    if isinstance(obj, (float, numpy.floating)):
        _check(lib.dasp_cent_update_float(ptr, code, float(obj)))
    elif isinstance(obj, (int, numpy.integer)):
        _check(lib.dasp_cent_update_int(ptr, code, int(obj)))
    else:
        raise error("centmodule: cannot use a %s for update code %d" % (type(obj).__name__, code))
    return None


def free(handle):
    """cent.free (cmod/src/centmodule.c:3163-3186)."""
    h = _handles.pop(handle, None)
    if h is not None:
        _check(library().dasp_cent_free(h.ptr))
    return None


def testcorrelation(img, corr):
    """cent.testcorrelation (cmod/src/centmodule.c:3275-3309): corr <- correlation(img, corr)."""
    if not _is_f32(img) or not _is_f32(corr):
        raise error("Error: centmodule - img must be float and contiguous")
    if img.ndim != 2 or corr.ndim != 2 or img.shape[0] != img.shape[1] or corr.shape != img.shape:
        raise error("Error: centmodule - corr/img wrong shape")
    _check(library().dasp_cent_testcorrelation(img.ctypes.data, corr.ctypes.data, img.shape[0]))
    return None


# ---- extensions ---------------------------------------------------------------------------------
FORM_FFT32, FORM_DFT_BF16, FORM_DFT_FP16, FORM_FFT32_GENERAL = 0, 1, 2, 3


def setForm(handle, form):
    """Choose the compute form: FORM_FFT32 (default; fused kernel where the options allow),
    FORM_DFT_BF16, FORM_DFT_FP16, or FORM_FFT32_GENERAL (FP32, general kernels only)."""
    _check(library().dasp_cent_set_form(_ptr(handle), int(form)))


def setFastTrig(handle, enable=True):
    """Use the reduced-argument hardware sine/cosine for exp(i(phi - t)) in the fused kernels."""
    _check(library().dasp_cent_set_fast_trig(_ptr(handle), int(bool(enable))))


def info(handle):
    """(nactive, npup, corrsize) of a handle."""
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(library().dasp_cent_info(_ptr(handle), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
    return a.value, b.value, c.value


def runDevice(handle, phs_ptr, phs_stride, nframes, cents_ptr, bimg_ptr=0, corrimg_ptr=0,
              zpois_ptr=0, zread_ptr=0, stream=0):
    """Device-resident step over nframes frames (all arguments are device addresses)."""
    _check(library().dasp_cent_run_device(_ptr(handle), phs_ptr, int(phs_stride), int(nframes), cents_ptr,
                                          bimg_ptr or None, corrimg_ptr or None, zpois_ptr or None,
                                          zread_ptr or None, stream or None))


def drawNormals(seed, stream, n):
    """The engine's host noise stream: n float32 N(0,1) values (PCG32 + Box-Muller)."""
    out = numpy.empty(int(n), numpy.float32)
    _check(library().dasp_cent_draw_normals(int(seed), int(stream), int(n), out.ctypes.data))
    return out


def drawNormalsDevice(seed, stream, n, out_ptr, hip_stream=0):
    """The same stream written by the device into n floats at device address out_ptr,
    asynchronously on hip_stream (values equal drawNormals' up to the last float bit)."""
    _check(library().dasp_cent_draw_normals_device(int(seed), int(stream), int(n), int(out_ptr),
                                                   int(hip_stream)))


def hcToSpectrum(pattern_hc):
    """Complex spectrum equivalent to one half-complex correlation pattern (C, C)."""
    p = numpy.ascontiguousarray(pattern_hc, numpy.float32)
    C = p.shape[0]
    out = numpy.empty((C, C, 2), numpy.float32)
    _check(library().dasp_cent_hc_to_spectrum(p.ctypes.data, C, out.ctypes.data))
    return out[..., 0] + 1j * out[..., 1]


def setTiming(handle, enable=True):
    """Bracket every later launch of the main image kernel with HIP events on its stream."""
    _check(library().dasp_cent_set_timing(_ptr(handle), int(bool(enable))))


def kernelTime(handle, reset=False):
    """(total milliseconds, launches) of the main image kernel since the last reset."""
    ms, n = ctypes.c_double(), ctypes.c_long()
    _check(library().dasp_cent_kernel_time(_ptr(handle), int(bool(reset)), ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value


# ---- multi-WFS step ------------------------------------------------------------------------
def batchCreate(items):
    """Build a multi-WFS step. items: sequence of dicts with keys handle, phs, phs_stride, nframes,
    cents and optionally bimg, corrimg, zpois, zread (device addresses). Returns a batch handle."""
    arr = (_BatchItem * len(items))()
    for k, it in enumerate(items):
        arr[k].handle = _ptr(it["handle"])
        arr[k].phs = it["phs"]
        arr[k].phs_stride = int(it["phs_stride"])
        arr[k].nframes = int(it["nframes"])
        arr[k].cents = it["cents"]
        for key in ("bimg", "corrimg", "zpois", "zread"):
            setattr(arr[k], key, it.get(key) or None)
    out = ctypes.c_void_p()
    _check(library().dasp_cent_batch_create(arr, len(items), ctypes.byref(out)))
    return out.value


def batchTune(batch, stream=0, reps=5):
    """Autotune waves per workgroup and workgroup count of every persistent launch (synchronous)."""
    _check(library().dasp_cent_batch_tune(batch, stream or None, int(reps)))


def batchRun(batch, stream=0):
    """One step: a single graph launch on stream (asynchronous)."""
    _check(library().dasp_cent_batch_run(batch, stream or None))


def batchInfo(batch):
    """List of dicts describing the persistent launches, and the number of WFS outside them."""
    lib = library()
    n = lib.dasp_cent_batch_launches(batch)
    if n < 0:
        _check(n)
    out = []
    for i in range(n):
        shape = (ctypes.c_int * 5)()
        nwfs, waves, chunk = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        groups, wgs = ctypes.c_long(), ctypes.c_long()
        ms = ctypes.c_float()
        _check(lib.dasp_cent_batch_launch_info(batch, i, shape, ctypes.byref(nwfs), ctypes.byref(groups),
                                               ctypes.byref(waves), ctypes.byref(wgs), ctypes.byref(chunk),
                                               ctypes.byref(ms)))
        out.append(dict(np=shape[0], fftsize=shape[1], nimg=shape[2], spot=bool(shape[3]), fast=bool(shape[4]),
                        nwfs=nwfs.value, groups=groups.value,
                        waves_per_wg=waves.value, workgroups=wgs.value, chunk=chunk.value, tuned_ms=ms.value))
    others = lib.dasp_cent_batch_others(batch)
    if others < 0:
        _check(others)
    return out, others


def batchSetConfig(batch, i, waves_per_wg, workgroups, chunk=0):
    """Apply a launch configuration (as reported by batchInfo) and rebuild the step graph."""
    _check(library().dasp_cent_batch_set_config(batch, int(i), int(waves_per_wg), int(workgroups), int(chunk)))


def batchFree(batch):
    _check(library().dasp_cent_batch_free(batch))
