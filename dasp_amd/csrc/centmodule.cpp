// centmodule.cpp - the CPython extension module "cent" (PyInit_cent): the drop-in for DASP's
// cmod.cent (cmod/src/centmodule.c:3313-3332), built like the reference's cmod extensions
// (setup.py Extension with the engine's hipcc objects as extra_objects, cmod/src/setup.py:55-63,
// :76-82). Every function converts Python arguments and calls the C ABI of include/dasp_cent.h.
//
// Functions of the reference, same names, argument order and behaviour:
//   initialise(41 args) -> int handle     (py_initialise :2858-3161, format string :2876)
//   run(handle) -> float seconds           (py_run :3188-3229)
//   runSlope(handle) -> float seconds      (py_runSlope :3230-3272)
//   update(handle, code, value)            (py_update :2573-2856)
//   free(handle)                           (py_free :3163-3186)
//   testcorrelation(img, corr)             (py_testcorrelation :3275-3309)
// As in the reference the numpy arrays are borrowed (their data pointers are kept); this module
// also holds a reference to each of them until free(), so they cannot be collected under it.
// Failures raise cent.error with the engine's message (the reference printed and returned NULL).
//
// Extensions (no counterpart in the reference): setForm, setFastTrig, setRelativeFloor, setGrid,
// getGrid, tuneGrid, info,
// runDevice, runDeviceNoAlloc, drawNormals, drawNormalsDevice, hcToSpectrum, setTiming, kernelTime, abiVersion and
// the multi-WFS step batchCreate / batchTune / batchRun / batchInfo / batchSetConfig / batchFree,
// the device reconstructor step reconCreate / reconSetGains / reconRun / reconFree, and the
// device phase producer (include/dasp_atmos.h) atmosPupilPhase / atmosLayerUpdate /
// atmosWorkDoubles / dmSplineOperator / dmSurface / dmLineOfSight.
#define PY_SSIZE_T_CLEAN
#include <Python.h>
#define NPY_NO_DEPRECATED_API NPY_1_7_API_VERSION
#include <numpy/arrayobject.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/dasp_atmos.h"
#include "../../include/dasp_cent.h"

static_assert(DASP_CENT_ABI_VERSION == 2, "module written for ABI version 2");
static_assert(sizeof(dasp_cent_params) == 856 && sizeof(dasp_cent_batch_item) == 72, "ABI layout changed");

static PyObject *CentError = nullptr;
static PyObject *Kept = nullptr;   // dict: handle -> list of the borrowed Python objects

static PyObject *engine_error(int rc) {
  const char *m = dasp_cent_last_error();
  if (m && m[0]) PyErr_SetString(CentError, m);
  else PyErr_Format(CentError, "error %d", rc);
  return nullptr;
}

#define CHECK(call)                         \
  do {                                      \
    int _rc = (call);                       \
    if (_rc != 0) return engine_error(_rc); \
  } while (0)

// A dasp_cent_array over a numpy buffer (checkFloatContigArr / checkIntContigArr, :2132-2149).
// kind 'f': float32, 'i': 4-byte signed integer. A non-array object stands for "not given" when
// allow_none (the reference tests PyArray_Check and otherwise leaves the pointer NULL).
static int to_array(PyObject *o, dasp_cent_array *a, char kind, const char *what, bool allow_none) {
  std::memset(a, 0, sizeof(*a));
  if (o == nullptr || o == Py_None || !PyArray_Check(o)) {
    if (allow_none) return 0;
    PyErr_Format(CentError, "centmodule - %s must be a%s array", what, kind == 'f' ? " float32" : "n int32");
    return -1;
  }
  PyArrayObject *arr = (PyArrayObject *)o;
  const int t = PyArray_TYPE(arr);
  const bool ok_t = kind == 'f' ? t == NPY_FLOAT32 : (PyArray_ITEMSIZE(arr) == 4 && PyArray_ISSIGNED(arr) && PyArray_ISINTEGER(arr));
  if (!ok_t || !PyArray_IS_C_CONTIGUOUS(arr)) {
    PyErr_Format(CentError, "centmodule - %s must be %s and contiguous", what, kind == 'f' ? "float32" : "int32");
    return -1;
  }
  if (PyArray_NDIM(arr) > 4) {
    PyErr_Format(CentError, "centmodule - %s has too many dimensions", what);
    return -1;
  }
  a->data = PyArray_DATA(arr);
  a->ndim = PyArray_NDIM(arr);
  for (int i = 0; i < a->ndim; ++i) a->dims[i] = (long)PyArray_DIM(arr, i);
  return 0;
}

static bool to_handle(PyObject *o, dasp_cent_t **h) {
  void *p = PyLong_AsVoidPtr(o);
  if (PyErr_Occurred()) return false;
  PyObject *key = PyLong_FromVoidPtr(p);
  const int known = key ? PyDict_Contains(Kept, key) : -1;
  Py_XDECREF(key);
  if (known != 1) {
    if (known == 0) PyErr_Format(CentError, "invalid centstruct handle");
    return false;
  }
  *h = (dasp_cent_t *)p;
  return true;
}

static int keep(dasp_cent_t *h, PyObject *o) {
  PyObject *key = PyLong_FromVoidPtr(h);
  if (!key) return -1;
  PyObject *lst = PyDict_GetItem(Kept, key);
  Py_DECREF(key);
  return lst ? PyList_Append(lst, o) : 0;
}

// ---- reference functions ----------------------------------------------------------------
static PyObject *py_initialise(PyObject *, PyObject *args) {
  dasp_cent_params p;
  std::memset(&p, 0, sizeof(p));
  PyObject *sig, *sky, *phs, *pup, *spot, *cents, *subflag, *bimg, *frac, *cw, *cpat, *cimg, *ub, *inimg, *loc;
  long seed = 0;
  // the reference's format "iiiiiiiffifOOifiilOO!OO!O!O!O!iOifOOiiOiiiffOO" (:2876) with the typed
  // arrays checked here so that a wrong type raises cent.error like every other check
  if (!PyArg_ParseTuple(args, "iiiiiiiffifOOifiilOOOOOOOiOifOOiiOiiiffOO", &p.nthreads, &p.nsubaps, &p.ncen,
                        &p.fftsize, &p.clipsize, &p.nimg, &p.phasesize, &p.readnoise, &p.readbg, &p.addPoisson,
                        &p.noiseFloor, &sig, &sky, &p.calsource, &p.pxlPower, &p.nintegrations, &p.nlatency, &seed,
                        &phs, &pup, &spot, &cents, &subflag, &bimg, &frac, &p.opticalBinning, &cw,
                        &p.correlationCentroiding, &p.corrThresh, &cpat, &cimg, &p.threshType, &p.imageOnly, &ub,
                        &p.preBinningFactor, &p.parabolicFit, &p.gaussianFit, &p.gaussianMinVal,
                        &p.gaussianReplaceVal, &inimg, &loc))
    return nullptr;
  p.seed = (unsigned long)seed;
  if (PyArray_Check(sig)) {
    if (to_array(sig, &p.sigArr, 'f', "sig array", false)) return nullptr;
  } else if (PyFloat_Check(sig) || PyLong_Check(sig) || PyArray_IsScalar(sig, Floating)) {
    p.sig = (float)PyFloat_AsDouble(sig);
  } else {
    PyErr_SetString(CentError, "Error: centmodule - sig object must be float or float array");
    return nullptr;
  }
  if (PyArray_Check(sky)) {
    if (to_array(sky, &p.skybrightnessArr, 'f', "skybrightness", false)) return nullptr;
  } else {
    p.skybrightness = (float)PyFloat_AsDouble(sky);
    if (PyErr_Occurred()) return nullptr;
  }
  if (PyArray_Check(ub)) {
    if (to_array(ub, &p.useBrightestArr, 'i', "useBrightest", false)) return nullptr;
  } else if (PyLong_Check(ub) || PyArray_IsScalar(ub, Integer)) {
    p.useBrightest = (int)PyLong_AsLong(ub);
  } else {
    PyErr_SetString(CentError, "Error: centmodule - useBrightest must be int or int array");
    return nullptr;
  }
  if (to_array(phs, &p.phs, 'f', "phs", true) || to_array(pup, &p.pupfn, 'f', "pupfn", false) ||
      to_array(spot, &p.spotpsf, 'f', "spotpsf", true) || to_array(cents, &p.cents, 'f', "cents", false) ||
      to_array(subflag, &p.subflag, 'i', "subflag", false) || to_array(bimg, &p.bimg, 'f', "bimg", false) ||
      to_array(frac, &p.fracSubArea, 'f', "fracSubArea", false) || to_array(cw, &p.centWeight, 'f', "centWeight", true) ||
      to_array(cpat, &p.corrPattern, 'f', "corrPattern", true) || to_array(cimg, &p.corrimg, 'f', "corrimg", true) ||
      to_array(inimg, &p.inputImage, 'f', "inputImage", true) ||
      to_array(loc, &p.subapLocation, 'i', "subapLocation", true))
    return nullptr;
  dasp_cent_t *h = nullptr;
  int rc;
  Py_BEGIN_ALLOW_THREADS
    rc = dasp_cent_initialise(&p, &h);
  Py_END_ALLOW_THREADS
  if (rc != 0) return engine_error(rc);
  PyObject *key = PyLong_FromVoidPtr(h);
  PyObject *lst = Py_BuildValue("[OOOOOOOOOOOOOOO]", phs, pup, spot, cents, subflag, bimg, frac, cw, cpat, cimg,
                                inimg, loc, sig, sky, ub);
  if (!key || !lst || PyDict_SetItem(Kept, key, lst) != 0) {
    Py_XDECREF(key);
    Py_XDECREF(lst);
    dasp_cent_free(h);
    return nullptr;
  }
  Py_DECREF(lst);
  return key;
}

static PyObject *py_run(PyObject *, PyObject *args) {
  PyObject *ho;
  dasp_cent_t *h;
  if (!PyArg_ParseTuple(args, "O", &ho) || !to_handle(ho, &h)) return nullptr;
  float t = 0.f;
  int rc;
  Py_BEGIN_ALLOW_THREADS
    rc = dasp_cent_run(h, &t);
  Py_END_ALLOW_THREADS
  if (rc != 0) return engine_error(rc);
  return PyFloat_FromDouble(t);
}

static PyObject *py_runSlope(PyObject *, PyObject *args) {
  PyObject *ho;
  dasp_cent_t *h;
  if (!PyArg_ParseTuple(args, "O", &ho) || !to_handle(ho, &h)) return nullptr;
  float t = 0.f;
  int rc;
  Py_BEGIN_ALLOW_THREADS
    rc = dasp_cent_run_slope(h, &t);
  Py_END_ALLOW_THREADS
  if (rc != 0) return engine_error(rc);
  return PyFloat_FromDouble(t);
}

static bool is_array_code(int code) {
  return code == DASP_CENT_SKYBRIGHTNESS || code == DASP_CENT_SPOTPSF || code == DASP_CENT_CENTWEIGHT ||
         code == DASP_CENT_CORRPATTERN || code == DASP_CENT_REFCENTS || code == DASP_CENT_CALCOEFF ||
         code == DASP_CENT_OUTPUTPHASEARR;
}

// py_update dispatches on the code, then on the Python type of the value (:2573-2856).
static PyObject *py_update(PyObject *, PyObject *args) {
  PyObject *ho, *obj;
  int code;
  dasp_cent_t *h;
  if (!PyArg_ParseTuple(args, "OiO", &ho, &code, &obj) || !to_handle(ho, &h)) return nullptr;
  if (code == DASP_CENT_CALDATA) {   // None or (calData, calBounds, calSteps), :2153-2203
    if (obj == Py_None) {
      CHECK(dasp_cent_update_caldata(h, nullptr, nullptr, nullptr));
      Py_RETURN_NONE;
    }
    if (!PyTuple_Check(obj) || PyTuple_Size(obj) < 3) {
      PyErr_SetString(CentError, "Error: centmodule - calibration stuff not a tuple");
      return nullptr;
    }
    dasp_cent_array cd, cb, cs;
    if (to_array(PyTuple_GET_ITEM(obj, 0), &cd, 'f', "calData", false) ||
        to_array(PyTuple_GET_ITEM(obj, 1), &cb, 'i', "calBounds", false) ||
        to_array(PyTuple_GET_ITEM(obj, 2), &cs, 'f', "calSteps", false))
      return nullptr;
    CHECK(dasp_cent_update_caldata(h, &cd, &cb, &cs));
    if (keep(h, obj)) return nullptr;
    Py_RETURN_NONE;
  }
  if (PyArray_Check(obj) && PyArray_NDIM((PyArrayObject *)obj) > 0) {
    dasp_cent_array a;
    if (to_array(obj, &a, code == DASP_CENT_USEBRIGHTEST ? 'i' : 'f', "value", false)) return nullptr;
    CHECK(dasp_cent_update_array(h, code, &a));
    if (keep(h, obj)) return nullptr;
    Py_RETURN_NONE;
  }
  const bool number = PyFloat_Check(obj) || PyLong_Check(obj) || PyArray_IsScalar(obj, Number) || PyArray_Check(obj);
  if (obj == Py_None || (is_array_code(code) && !number)) {
    dasp_cent_array a;
    std::memset(&a, 0, sizeof(a));
    CHECK(dasp_cent_update_array(h, code, &a));
    Py_RETURN_NONE;
  }
  if (PyFloat_Check(obj) || PyArray_IsScalar(obj, Floating) ||
      (PyArray_Check(obj) && PyArray_ISFLOAT((PyArrayObject *)obj))) {
    const double v = PyFloat_AsDouble(obj);
    if (PyErr_Occurred()) return nullptr;
    CHECK(dasp_cent_update_float(h, code, v));
  } else if (PyLong_Check(obj) || PyArray_IsScalar(obj, Integer) ||
             (PyArray_Check(obj) && PyArray_ISINTEGER((PyArrayObject *)obj))) {
    const long v = PyLong_AsLong(obj);
    if (PyErr_Occurred()) return nullptr;
    CHECK(dasp_cent_update_int(h, code, v));
  } else {
    PyErr_Format(CentError, "centmodule: cannot use a %s for update code %d", Py_TYPE(obj)->tp_name, code);
    return nullptr;
  }
  Py_RETURN_NONE;
}

static PyObject *py_free(PyObject *, PyObject *args) {
  PyObject *ho;
  if (!PyArg_ParseTuple(args, "O", &ho)) return nullptr;
  void *p = PyLong_AsVoidPtr(ho);
  if (PyErr_Occurred()) return nullptr;
  PyObject *key = PyLong_FromVoidPtr(p);
  if (!key) return nullptr;
  const int known = PyDict_Contains(Kept, key);
  if (known == 1) {
    const int rc = dasp_cent_free((dasp_cent_t *)p);
    PyDict_DelItem(Kept, key);
    Py_DECREF(key);
    if (rc != 0) return engine_error(rc);
    Py_RETURN_NONE;
  }
  Py_DECREF(key);
  if (known < 0) return nullptr;
  Py_RETURN_NONE;   // freeing twice is harmless
}

static PyObject *py_testcorrelation(PyObject *, PyObject *args) {
  PyObject *io, *co;
  if (!PyArg_ParseTuple(args, "OO", &io, &co)) return nullptr;
  dasp_cent_array img, corr;
  if (to_array(io, &img, 'f', "img", false) || to_array(co, &corr, 'f', "corr", false)) return nullptr;
  if (img.ndim != 2 || corr.ndim != 2 || img.dims[0] != img.dims[1] || corr.dims[0] != img.dims[0] ||
      corr.dims[1] != img.dims[1]) {
    PyErr_SetString(CentError, "Error: centmodule - corr/img wrong shape");
    return nullptr;
  }
  CHECK(dasp_cent_testcorrelation((const float *)img.data, (float *)corr.data, (int)img.dims[0]));
  Py_RETURN_NONE;
}

// ---- extensions -------------------------------------------------------------------------
static PyObject *handle_int(PyObject *args, PyObject *kw, int (*fn)(dasp_cent_t *, int), int dflt) {
  static const char *names[] = {"handle", "enable", nullptr};
  PyObject *ho;
  int v = dflt;
  dasp_cent_t *h;
  if (!PyArg_ParseTupleAndKeywords(args, kw, "O|p", (char **)names, &ho, &v) || !to_handle(ho, &h)) return nullptr;
  CHECK(fn(h, v));
  Py_RETURN_NONE;
}

static PyObject *py_setForm(PyObject *, PyObject *args) {
  PyObject *ho;
  int form;
  dasp_cent_t *h;
  if (!PyArg_ParseTuple(args, "Oi", &ho, &form) || !to_handle(ho, &h)) return nullptr;
  CHECK(dasp_cent_set_form(h, form));
  Py_RETURN_NONE;
}
static PyObject *py_setFastTrig(PyObject *, PyObject *args, PyObject *kw) {
  return handle_int(args, kw, dasp_cent_set_fast_trig, 1);
}
static PyObject *py_setRelativeFloor(PyObject *, PyObject *args, PyObject *kw) {
  return handle_int(args, kw, dasp_cent_set_relative_floor, 1);
}
static PyObject *py_setTiming(PyObject *, PyObject *args, PyObject *kw) { return handle_int(args, kw, dasp_cent_set_timing, 1); }

static PyObject *py_info(PyObject *, PyObject *args) {
  PyObject *ho;
  dasp_cent_t *h;
  if (!PyArg_ParseTuple(args, "O", &ho) || !to_handle(ho, &h)) return nullptr;
  int a = 0, b = 0, c = 0;
  CHECK(dasp_cent_info(h, &a, &b, &c));
  return Py_BuildValue("(iii)", a, b, c);
}

static PyObject *py_kernelTime(PyObject *, PyObject *args, PyObject *kw) {
  static const char *names[] = {"handle", "reset", nullptr};
  PyObject *ho;
  int reset = 0;
  dasp_cent_t *h;
  if (!PyArg_ParseTupleAndKeywords(args, kw, "O|p", (char **)names, &ho, &reset) || !to_handle(ho, &h)) return nullptr;
  double ms = 0.0;
  long n = 0;
  CHECK(dasp_cent_kernel_time(h, reset, &ms, &n));
  return Py_BuildValue("(dl)", ms, n);
}

static void *as_ptr(unsigned long long v) { return (void *)(uintptr_t)v; }

static PyObject *py_runDevice(PyObject *, PyObject *args, PyObject *kw) {
  static const char *names[] = {"handle", "phs_ptr", "phs_stride", "nframes", "cents_ptr", "bimg_ptr", "corrimg_ptr",
                                "zpois_ptr", "zread_ptr", "stream", nullptr};
  PyObject *ho;
  unsigned long long phs, cents, bimg = 0, corr = 0, zp = 0, zr = 0, st = 0;
  long stride;
  int nframes;
  dasp_cent_t *h;
  if (!PyArg_ParseTupleAndKeywords(args, kw, "OKliK|KKKKK", (char **)names, &ho, &phs, &stride, &nframes, &cents,
                                   &bimg, &corr, &zp, &zr, &st) ||
      !to_handle(ho, &h))
    return nullptr;
  int rc;
  Py_BEGIN_ALLOW_THREADS
    rc = dasp_cent_run_device(h, (const float *)as_ptr(phs), stride, nframes, (float *)as_ptr(cents),
                                                   (float *)as_ptr(bimg), (float *)as_ptr(corr),
                                                   (const float *)as_ptr(zp), (const float *)as_ptr(zr), as_ptr(st));
  Py_END_ALLOW_THREADS
  if (rc != 0) return engine_error(rc);
  Py_RETURN_NONE;
}

static PyObject *py_setGrid(PyObject *, PyObject *args, PyObject *kw) {
  static const char *names[] = {"handle", "image", "corr", nullptr};
  PyObject *ho;
  long gi = 0, gc = 0;
  dasp_cent_t *h;
  if (!PyArg_ParseTupleAndKeywords(args, kw, "O|ll", (char **)names, &ho, &gi, &gc) || !to_handle(ho, &h)) return nullptr;
  const int rc = dasp_cent_set_grid(h, gi, gc);
  if (rc != 0) return engine_error(rc);
  Py_RETURN_NONE;
}

static PyObject *py_getGrid(PyObject *, PyObject *args) {
  PyObject *ho;
  dasp_cent_t *h;
  if (!PyArg_ParseTuple(args, "O", &ho) || !to_handle(ho, &h)) return nullptr;
  long gi = 0, gc = 0;
  const int rc = dasp_cent_get_grid(h, &gi, &gc);
  if (rc != 0) return engine_error(rc);
  return Py_BuildValue("(ll)", gi, gc);
}

static PyObject *py_tuneGrid(PyObject *, PyObject *args, PyObject *kw) {
  static const char *names[] = {"handle", "phs_ptr", "phs_stride", "nframes", "cents_ptr", "zpois_ptr", "zread_ptr",
                                "reps", "stream", nullptr};
  PyObject *ho;
  unsigned long long phs, cents, zp = 0, zr = 0, st = 0;
  long stride;
  int nframes, reps = 5;
  dasp_cent_t *h;
  if (!PyArg_ParseTupleAndKeywords(args, kw, "OKliK|KKiK", (char **)names, &ho, &phs, &stride, &nframes, &cents, &zp,
                                   &zr, &reps, &st) ||
      !to_handle(ho, &h))
    return nullptr;
  int rc;
  double ms = 0.0;
  Py_BEGIN_ALLOW_THREADS
    rc = dasp_cent_tune_grid(h, (const float *)as_ptr(phs), stride, nframes, (float *)as_ptr(cents),
                             (const float *)as_ptr(zp), (const float *)as_ptr(zr), reps, as_ptr(st), &ms);
  Py_END_ALLOW_THREADS
  if (rc != 0) return engine_error(rc);
  long gi = 0, gc = 0;
  (void)dasp_cent_get_grid(h, &gi, &gc);
  return Py_BuildValue("(lld)", gi, gc, ms);
}

static PyObject *py_runDeviceNoAlloc(PyObject *, PyObject *args, PyObject *kw) {
  static const char *names[] = {"handle", "phs_ptr", "phs_stride", "nframes", "cents_ptr", "bimg_ptr", "corrimg_ptr",
                                "zpois_ptr", "zread_ptr", "stream", nullptr};
  PyObject *ho;
  unsigned long long phs, cents, bimg = 0, corr = 0, zp = 0, zr = 0, st = 0;
  long stride;
  int nframes;
  dasp_cent_t *h;
  if (!PyArg_ParseTupleAndKeywords(args, kw, "OKliK|KKKKK", (char **)names, &ho, &phs, &stride, &nframes, &cents,
                                   &bimg, &corr, &zp, &zr, &st) ||
      !to_handle(ho, &h))
    return nullptr;
  int rc;
  Py_BEGIN_ALLOW_THREADS
    rc = dasp_cent_run_device_noalloc(h, (const float *)as_ptr(phs), stride, nframes, (float *)as_ptr(cents),
                                      (float *)as_ptr(bimg), (float *)as_ptr(corr), (const float *)as_ptr(zp),
                                      (const float *)as_ptr(zr), as_ptr(st));
  Py_END_ALLOW_THREADS
  if (rc != 0) return engine_error(rc);
  Py_RETURN_NONE;
}

static PyObject *py_drawNormals(PyObject *, PyObject *args) {
  unsigned long long seed, stream;
  long n;
  if (!PyArg_ParseTuple(args, "KKl", &seed, &stream, &n)) return nullptr;
  if (n < 0) {
    PyErr_SetString(CentError, "drawNormals: negative size");
    return nullptr;
  }
  npy_intp d = n;
  PyObject *out = PyArray_SimpleNew(1, &d, NPY_FLOAT32);
  if (!out) return nullptr;
  const int rc = dasp_cent_draw_normals((unsigned long)seed, (unsigned long)stream, n,
                                        (float *)PyArray_DATA((PyArrayObject *)out));
  if (rc != 0) {
    Py_DECREF(out);
    return engine_error(rc);
  }
  return out;
}

static PyObject *py_drawNormalsDevice(PyObject *, PyObject *args, PyObject *kw) {
  static const char *names[] = {"seed", "stream", "n", "out_ptr", "hip_stream", nullptr};
  unsigned long long seed, stream, out, st = 0;
  long n;
  if (!PyArg_ParseTupleAndKeywords(args, kw, "KKlK|K", (char **)names, &seed, &stream, &n, &out, &st)) return nullptr;
  CHECK(dasp_cent_draw_normals_device((unsigned long)seed, (unsigned long)stream, n, (float *)as_ptr(out), as_ptr(st)));
  Py_RETURN_NONE;
}

static PyObject *py_hcToSpectrum(PyObject *, PyObject *args) {
  PyObject *po;
  if (!PyArg_ParseTuple(args, "O", &po)) return nullptr;
  PyObject *arr = PyArray_FROMANY(po, NPY_FLOAT32, 2, 2, NPY_ARRAY_C_CONTIGUOUS | NPY_ARRAY_ALIGNED);
  if (!arr) return nullptr;
  const int C = (int)PyArray_DIM((PyArrayObject *)arr, 0);
  if (PyArray_DIM((PyArrayObject *)arr, 1) != C) {
    Py_DECREF(arr);
    PyErr_SetString(CentError, "hcToSpectrum: the pattern must be square");
    return nullptr;
  }
  std::vector<float> sp((size_t)2 * C * C);
  const int rc = dasp_cent_hc_to_spectrum((const float *)PyArray_DATA((PyArrayObject *)arr), C, sp.data());
  Py_DECREF(arr);
  if (rc != 0) return engine_error(rc);
  npy_intp d[2] = {C, C};
  PyObject *out = PyArray_SimpleNew(2, d, NPY_COMPLEX128);
  if (!out) return nullptr;
  double *o = (double *)PyArray_DATA((PyArrayObject *)out);
  for (size_t q = 0; q < sp.size(); ++q) o[q] = sp[q];
  return out;
}

static PyObject *py_abiVersion(PyObject *, PyObject *) { return PyLong_FromLong(dasp_cent_abi_version()); }

// ---- multi-WFS step ---------------------------------------------------------------------
static bool dict_ptr(PyObject *d, const char *key, void **out, bool required) {
  PyObject *v = PyDict_GetItemString(d, key);
  if (!v || v == Py_None) {
    *out = nullptr;
    if (required) PyErr_Format(CentError, "batch item needs '%s'", key);
    return !required;
  }
  const unsigned long long u = PyLong_AsUnsignedLongLong(v);
  if (PyErr_Occurred()) return false;
  *out = as_ptr(u);
  return true;
}

static PyObject *py_batchCreate(PyObject *, PyObject *args) {
  PyObject *seq;
  if (!PyArg_ParseTuple(args, "O", &seq)) return nullptr;
  PyObject *fast = PySequence_Fast(seq, "batchCreate needs a sequence of dicts");
  if (!fast) return nullptr;
  const Py_ssize_t n = PySequence_Fast_GET_SIZE(fast);
  std::vector<dasp_cent_batch_item> items((size_t)n);
  for (Py_ssize_t k = 0; k < n; ++k) {
    PyObject *d = PySequence_Fast_GET_ITEM(fast, k);
    dasp_cent_batch_item &it = items[(size_t)k];
    std::memset(&it, 0, sizeof(it));
    PyObject *ho = PyDict_Check(d) ? PyDict_GetItemString(d, "handle") : nullptr;
    PyObject *so = PyDict_Check(d) ? PyDict_GetItemString(d, "phs_stride") : nullptr;
    PyObject *fo = PyDict_Check(d) ? PyDict_GetItemString(d, "nframes") : nullptr;
    void *phs, *cents, *bimg, *corr, *zp, *zr;
    if (!ho || !so || !fo) {
      Py_DECREF(fast);
      PyErr_SetString(CentError, "batch item needs handle, phs, phs_stride, nframes and cents");
      return nullptr;
    }
    if (!to_handle(ho, &it.handle) || !dict_ptr(d, "phs", &phs, true) || !dict_ptr(d, "cents", &cents, true) ||
        !dict_ptr(d, "bimg", &bimg, false) || !dict_ptr(d, "corrimg", &corr, false) ||
        !dict_ptr(d, "zpois", &zp, false) || !dict_ptr(d, "zread", &zr, false)) {
      Py_DECREF(fast);
      return nullptr;
    }
    it.phs = (const float *)phs;
    it.phs_stride = PyLong_AsLong(so);
    it.nframes = (int)PyLong_AsLong(fo);
    it.cents = (float *)cents;
    it.bimg = (float *)bimg;
    it.corrimg = (float *)corr;
    it.zpois = (const float *)zp;
    it.zread = (const float *)zr;
    if (PyErr_Occurred()) {
      Py_DECREF(fast);
      return nullptr;
    }
  }
  Py_DECREF(fast);
  dasp_cent_batch_t *b = nullptr;
  CHECK(dasp_cent_batch_create(items.data(), (int)n, &b));
  return PyLong_FromVoidPtr(b);
}

static bool to_batch(PyObject *o, dasp_cent_batch_t **b) {
  *b = (dasp_cent_batch_t *)PyLong_AsVoidPtr(o);
  return !PyErr_Occurred();
}

static PyObject *py_batchTune(PyObject *, PyObject *args, PyObject *kw) {
  static const char *names[] = {"batch", "stream", "reps", nullptr};
  PyObject *bo;
  unsigned long long st = 0;
  int reps = 5;
  dasp_cent_batch_t *b;
  if (!PyArg_ParseTupleAndKeywords(args, kw, "O|Ki", (char **)names, &bo, &st, &reps) || !to_batch(bo, &b)) return nullptr;
  int rc;
  Py_BEGIN_ALLOW_THREADS
    rc = dasp_cent_batch_tune(b, as_ptr(st), reps);
  Py_END_ALLOW_THREADS
  if (rc != 0) return engine_error(rc);
  Py_RETURN_NONE;
}

static PyObject *py_batchRun(PyObject *, PyObject *args, PyObject *kw) {
  static const char *names[] = {"batch", "stream", nullptr};
  PyObject *bo;
  unsigned long long st = 0;
  dasp_cent_batch_t *b;
  if (!PyArg_ParseTupleAndKeywords(args, kw, "O|K", (char **)names, &bo, &st) || !to_batch(bo, &b)) return nullptr;
  CHECK(dasp_cent_batch_run(b, as_ptr(st)));
  Py_RETURN_NONE;
}

static PyObject *py_batchInfo(PyObject *, PyObject *args) {
  PyObject *bo;
  dasp_cent_batch_t *b;
  if (!PyArg_ParseTuple(args, "O", &bo) || !to_batch(bo, &b)) return nullptr;
  const int n = dasp_cent_batch_launches(b);
  if (n < 0) return engine_error(n);
  PyObject *lst = PyList_New(0);
  if (!lst) return nullptr;
  for (int i = 0; i < n; ++i) {
    int shape[5], nwfs, waves, chunk;
    long groups, wgs;
    float ms;
    const int rc = dasp_cent_batch_launch_info(b, i, shape, &nwfs, &groups, &waves, &wgs, &chunk, &ms);
    if (rc != 0) {
      Py_DECREF(lst);
      return engine_error(rc);
    }
    PyObject *d = Py_BuildValue("{s:i,s:i,s:i,s:O,s:O,s:i,s:l,s:i,s:l,s:i,s:d}", "np", shape[0], "fftsize", shape[1],
                                "nimg", shape[2], "spot", shape[3] ? Py_True : Py_False, "fast",
                                shape[4] ? Py_True : Py_False, "nwfs", nwfs, "groups", groups, "waves_per_wg", waves,
                                "workgroups", wgs, "chunk", chunk, "tuned_ms", (double)ms);
    if (!d || PyList_Append(lst, d) != 0) {
      Py_XDECREF(d);
      Py_DECREF(lst);
      return nullptr;
    }
    Py_DECREF(d);
  }
  const int others = dasp_cent_batch_others(b);
  if (others < 0) {
    Py_DECREF(lst);
    return engine_error(others);
  }
  return Py_BuildValue("(Ni)", lst, others);
}

static PyObject *py_batchSetConfig(PyObject *, PyObject *args, PyObject *kw) {
  static const char *names[] = {"batch", "i", "waves_per_wg", "workgroups", "chunk", nullptr};
  PyObject *bo;
  int i, waves, chunk = 0;
  long wgs;
  dasp_cent_batch_t *b;
  if (!PyArg_ParseTupleAndKeywords(args, kw, "Oiil|i", (char **)names, &bo, &i, &waves, &wgs, &chunk) || !to_batch(bo, &b))
    return nullptr;
  CHECK(dasp_cent_batch_set_config(b, i, waves, wgs, chunk));
  Py_RETURN_NONE;
}

static PyObject *py_batchFree(PyObject *, PyObject *args) {
  PyObject *bo;
  dasp_cent_batch_t *b;
  if (!PyArg_ParseTuple(args, "O", &bo) || !to_batch(bo, &b)) return nullptr;
  CHECK(dasp_cent_batch_free(b));
  Py_RETURN_NONE;
}

// ---- reconstructor step ---------------------------------------------------------------------
static PyObject *recon_error(int rc) {
  const char *m = dasp_recon_last_error();
  PyErr_SetString(CentError, (m && m[0]) ? m : "recon error");
  (void)rc;
  return nullptr;
}

static PyObject *py_reconCreate(PyObject *, PyObject *args) {
  int ncents, nacts, layout;
  unsigned long long mx;
  long ld;
  PyObject *io;
  if (!PyArg_ParseTuple(args, "iiiKlO", &ncents, &nacts, &layout, &mx, &ld, &io)) return nullptr;
  dasp_cent_array idx;
  if (to_array(io, &idx, 'i', "index", false)) return nullptr;
  if (idx.ndim != 1 || idx.dims[0] != ncents) {
    PyErr_SetString(CentError, "recon: index must be int32 of length ncents");
    return nullptr;
  }
  dasp_recon_t *r = nullptr;
  const int rc = dasp_recon_create(ncents, nacts, layout, (const float *)as_ptr(mx), ld, (const int *)idx.data, &r);
  if (rc != 0) return recon_error(rc);
  return PyLong_FromVoidPtr(r);
}

// gains / decay: a float for every actuator, or a float32 array of nacts
static PyObject *py_reconSetGains(PyObject *, PyObject *args) {
  PyObject *ro, *go, *dobj;
  if (!PyArg_ParseTuple(args, "OOO", &ro, &go, &dobj)) return nullptr;
  dasp_recon_t *r = (dasp_recon_t *)PyLong_AsVoidPtr(ro);
  if (PyErr_Occurred()) return nullptr;
  dasp_cent_array ga, da;
  float g = 1.f, d = 1.f;
  std::memset(&ga, 0, sizeof(ga));
  std::memset(&da, 0, sizeof(da));
  if (PyArray_Check(go)) {
    if (to_array(go, &ga, 'f', "gains", false)) return nullptr;
  } else {
    g = (float)PyFloat_AsDouble(go);
    if (PyErr_Occurred()) return nullptr;
  }
  if (PyArray_Check(dobj)) {
    if (to_array(dobj, &da, 'f', "decay", false)) return nullptr;
  } else {
    d = (float)PyFloat_AsDouble(dobj);
    if (PyErr_Occurred()) return nullptr;
  }
  const int rc = dasp_recon_set_gains(r, (const float *)ga.data, g, (const float *)da.data, d);
  if (rc != 0) return recon_error(rc);
  Py_RETURN_NONE;
}

static PyObject *py_reconRun(PyObject *, PyObject *args, PyObject *kw) {
  static const char *names[] = {"recon", "slopes_ptr", "slope_stride", "nframes", "acts_ptr", "acts_stride", "stream", nullptr};
  PyObject *ro;
  unsigned long long sl, ac, st = 0;
  long ss, as;
  int nframes;
  if (!PyArg_ParseTupleAndKeywords(args, kw, "OKliKl|K", (char **)names, &ro, &sl, &ss, &nframes, &ac, &as, &st)) return nullptr;
  dasp_recon_t *r = (dasp_recon_t *)PyLong_AsVoidPtr(ro);
  if (PyErr_Occurred()) return nullptr;
  const int rc = dasp_recon_run(r, (const float *)as_ptr(sl), ss, nframes, (float *)as_ptr(ac), as, as_ptr(st));
  if (rc != 0) return recon_error(rc);
  Py_RETURN_NONE;
}

static PyObject *py_reconFree(PyObject *, PyObject *args) {
  PyObject *ro;
  if (!PyArg_ParseTuple(args, "O", &ro)) return nullptr;
  dasp_recon_t *r = (dasp_recon_t *)PyLong_AsVoidPtr(ro);
  if (PyErr_Occurred()) return nullptr;
  dasp_recon_free(r);
  Py_RETURN_NONE;
}

// ---- device phase producer (atmosphere, deformable mirror) ------------------------------------
static PyObject *atmos_error(int rc) {
  const char *m = dasp_atmos_last_error();
  if (m && m[0]) PyErr_SetString(CentError, m);
  else PyErr_Format(CentError, "atmosphere/DM error %d", rc);
  return nullptr;
}

static PyObject *py_atmosWorkDoubles(PyObject *, PyObject *args) {
  int npup;
  if (!PyArg_ParseTuple(args, "i", &npup)) return nullptr;
  return PyLong_FromLong(dasp_atmos_work_doubles(npup));
}

// layers: sequence of tuples (screen_ptr, rows, cols, oy, ox, r0, c0, s1, s2, s3, s4)
static PyObject *py_atmosPupilPhase(PyObject *, PyObject *args, PyObject *kw) {
  static const char *names[] = {"layers", "npup", "intrinsic_ptr", "pupfn_ptr", "pupil_area", "remove_piston",
                                "out_scale", "out64_ptr", "out32_ptr", "work_ptr", "stream", nullptr};
  PyObject *seq;
  int npup, rp;
  unsigned long long intr, pup, o64, o32, work, st = 0;
  double area, scale;
  if (!PyArg_ParseTupleAndKeywords(args, kw, "OiKKdpdKKK|K", (char **)names, &seq, &npup, &intr, &pup, &area, &rp,
                                   &scale, &o64, &o32, &work, &st))
    return nullptr;
  PyObject *fast = PySequence_Fast(seq, "layers must be a sequence");
  if (!fast) return nullptr;
  const Py_ssize_t n = PySequence_Fast_GET_SIZE(fast);
  if (n > DASP_ATMOS_MAX_LAYERS) {
    Py_DECREF(fast);
    PyErr_Format(CentError, "at most %d layers", DASP_ATMOS_MAX_LAYERS);
    return nullptr;
  }
  std::vector<dasp_atmos_layer> L((size_t)n);
  for (Py_ssize_t k = 0; k < n; ++k) {
    unsigned long long sp;
    dasp_atmos_layer &y = L[(size_t)k];
    if (!PyArg_ParseTuple(PySequence_Fast_GET_ITEM(fast, k), "Kiiiiiiffff", &sp, &y.rows, &y.cols, &y.oy, &y.ox, &y.r0,
                          &y.c0, &y.s1, &y.s2, &y.s3, &y.s4)) {
      Py_DECREF(fast);
      return nullptr;
    }
    y.screen = (const double *)(uintptr_t)sp;
  }
  Py_DECREF(fast);
  const int rc = dasp_atmos_pupil_phase(L.data(), (int)n, npup, (const double *)(uintptr_t)intr, (const float *)(uintptr_t)pup,
                                        area, rp, scale, (double *)(uintptr_t)o64, (float *)(uintptr_t)o32,
                                        (double *)(uintptr_t)work, (void *)(uintptr_t)st);
  if (rc != 0) return atmos_error(rc);
  Py_RETURN_NONE;
}

static PyObject *py_atmosLayerUpdate(PyObject *, PyObject *args, PyObject *kw) {
  static const char *names[] = {"screen_ptr", "rows", "cols", "oy", "ox", "cols_ptr", "ncols", "rows_ptr", "nrows",
                                "col_add", "row_add", "stream", nullptr};
  unsigned long long scr, cp, rp, st = 0;
  int rows, cols, oy, ox, nc, nr, ca, ra;
  if (!PyArg_ParseTupleAndKeywords(args, kw, "KiiiiKiKiii|K", (char **)names, &scr, &rows, &cols, &oy, &ox, &cp, &nc,
                                   &rp, &nr, &ca, &ra, &st))
    return nullptr;
  const int rc = dasp_atmos_layer_update((double *)(uintptr_t)scr, rows, cols, &oy, &ox, (const double *)(uintptr_t)cp, nc,
                                         (const double *)(uintptr_t)rp, nr, ca, ra, (void *)(uintptr_t)st);
  if (rc != 0) return atmos_error(rc);
  return Py_BuildValue("ii", oy, ox);
}

static PyObject *py_dmSplineOperator(PyObject *, PyObject *args) {
  int npup, nact;
  double actoffset = 0.0;
  if (!PyArg_ParseTuple(args, "ii|d", &npup, &nact, &actoffset)) return nullptr;
  if (npup < 1 || nact < 2) {
    PyErr_SetString(CentError, "dmSplineOperator: npup >= 1 and nact >= 2 are required");
    return nullptr;
  }
  npy_intp dims[2] = {npup, nact};
  PyObject *op = PyArray_SimpleNew(2, dims, NPY_FLOAT64);
  if (!op) return nullptr;
  const int rc = dasp_dm_spline_operator(npup, nact, actoffset, (double *)PyArray_DATA((PyArrayObject *)op));
  if (rc != 0) {
    Py_DECREF(op);
    return atmos_error(rc);
  }
  return op;
}

static PyObject *py_dmSurface(PyObject *, PyObject *args, PyObject *kw) {
  static const char *names[] = {"act_ptr", "nact", "coupling", "op_ptr", "opT_ptr", "npup", "surface_ptr", "work_ptr",
                                "stream", nullptr};
  unsigned long long act, op, opT, surf, work, st = 0;
  int nact, npup;
  float coupling;
  if (!PyArg_ParseTupleAndKeywords(args, kw, "KifKKiKK|K", (char **)names, &act, &nact, &coupling, &op, &opT, &npup, &surf,
                                   &work, &st))
    return nullptr;
  const int rc = dasp_dm_surface((const float *)(uintptr_t)act, nact, coupling, (const double *)(uintptr_t)op,
                                 (const double *)(uintptr_t)opT, npup, (float *)(uintptr_t)surf, (void *)(uintptr_t)work,
                                 (void *)(uintptr_t)st);
  if (rc != 0) return atmos_error(rc);
  Py_RETURN_NONE;
}

static PyObject *py_dmLineOfSight(PyObject *, PyObject *args, PyObject *kw) {
  static const char *names[] = {"surface_ptr", "dmpup", "yoff", "xoff", "npup", "factor", "adjust", "base64_ptr",
                                "base32_ptr", "pupfn_ptr", "pupil_sum", "remove_piston", "out32_ptr", "work_ptr",
                                "stream", nullptr};
  unsigned long long surf, b64, b32, pup, out, work, st = 0;
  int dmpup, yoff, xoff, npup, adjust, rp;
  float factor;
  double psum;
  if (!PyArg_ParseTupleAndKeywords(args, kw, "KiiiifpKKKdpKK|K", (char **)names, &surf, &dmpup, &yoff, &xoff, &npup,
                                   &factor, &adjust, &b64, &b32, &pup, &psum, &rp, &out, &work, &st))
    return nullptr;
  const int rc = dasp_dm_line_of_sight((const float *)(uintptr_t)surf, dmpup, yoff, xoff, npup, factor, adjust,
                                       (const double *)(uintptr_t)b64, (const float *)(uintptr_t)b32,
                                       (const float *)(uintptr_t)pup, psum, rp, (float *)(uintptr_t)out,
                                       (double *)(uintptr_t)work, (void *)(uintptr_t)st);
  if (rc != 0) return atmos_error(rc);
  Py_RETURN_NONE;
}

static PyMethodDef centMethods[] = {
    {"run", py_run, METH_VARARGS, "Run the phase -> centroid algorithm; returns the elapsed seconds."},
    {"runSlope", py_runSlope, METH_VARARGS, "Run the image -> centroid algorithm."},
    {"initialise", py_initialise, METH_VARARGS, "Initialise the cent module (41 arguments); returns a handle."},
    {"update", py_update, METH_VARARGS, "update(handle, code, value): change a parameter."},
    {"free", py_free, METH_VARARGS, "Free a handle (invalidates it)."},
    {"testcorrelation", py_testcorrelation, METH_VARARGS, "testcorrelation(img, corr): corr <- correlation(img, corr)."},
    {"setForm", py_setForm, METH_VARARGS, "setForm(handle, form): FORM_FFT32, FORM_DFT_BF16, FORM_DFT_FP16 or FORM_FFT32_GENERAL."},
    {"setFastTrig", (PyCFunction)(void (*)(void))py_setFastTrig, METH_VARARGS | METH_KEYWORDS, "setFastTrig(handle, enable=True): reduced-argument hardware sin/cos."},
    {"setRelativeFloor", (PyCFunction)(void (*)(void))py_setRelativeFloor, METH_VARARGS | METH_KEYWORDS,
     "setRelativeFloor(handle, enable=True): threshType 0 floor = noiseFloor * max(image) when 0 < noiseFloor <= 1."},
    {"info", py_info, METH_VARARGS, "info(handle) -> (nactive, npup, corrsize)."},
    {"setGrid", (PyCFunction)(void (*)(void))py_setGrid, METH_VARARGS | METH_KEYWORDS,
     "setGrid(handle, image=0, corr=0): launch grids of the image and correlation kernels (0: default, <0: uncapped)."},
    {"getGrid", (PyCFunction)(void (*)(void))py_getGrid, METH_VARARGS, "getGrid(handle) -> (image_grid, corr_grid)."},
    {"tuneGrid", (PyCFunction)(void (*)(void))py_tuneGrid, METH_VARARGS | METH_KEYWORDS,
     "tuneGrid(handle, phs_ptr, phs_stride, nframes, cents_ptr, zpois_ptr=0, zread_ptr=0, reps=5, stream=0) -> "
     "(image_grid, corr_grid, ms): time the launch grids on these inputs and keep the fastest."},
    {"runDevice", (PyCFunction)(void (*)(void))py_runDevice, METH_VARARGS | METH_KEYWORDS,
     "runDevice(handle, phs_ptr, phs_stride, nframes, cents_ptr, bimg_ptr=0, corrimg_ptr=0, zpois_ptr=0, "
     "zread_ptr=0, stream=0): "
     "device-resident step over nframes frames (device addresses)."},
    {"runDeviceNoAlloc", (PyCFunction)(void (*)(void))py_runDeviceNoAlloc, METH_VARARGS | METH_KEYWORDS,
     "runDeviceNoAlloc(...): runDevice with its scratch in a fixed per-GPU array, no device allocation."},
    {"drawNormals", py_drawNormals, METH_VARARGS, "drawNormals(seed, stream, n) -> float32 N(0,1) array (PCG32 + Box-Muller)."},
    {"drawNormalsDevice", (PyCFunction)(void (*)(void))py_drawNormalsDevice, METH_VARARGS | METH_KEYWORDS,
     "drawNormalsDevice(seed, stream, n, out_ptr, hip_stream=0): the same stream written on the device."},
    {"hcToSpectrum", py_hcToSpectrum, METH_VARARGS, "hcToSpectrum(pattern_hc) -> complex spectrum of a half-complex pattern."},
    {"setTiming", (PyCFunction)(void (*)(void))py_setTiming, METH_VARARGS | METH_KEYWORDS, "setTiming(handle, enable=True): HIP events around the image kernel."},
    {"kernelTime", (PyCFunction)(void (*)(void))py_kernelTime, METH_VARARGS | METH_KEYWORDS, "kernelTime(handle, reset=False) -> (ms, launches)."},
    {"abiVersion", py_abiVersion, METH_NOARGS, "DASP_CENT_ABI_VERSION of the engine objects."},
    {"batchCreate", py_batchCreate, METH_VARARGS, "batchCreate(items) -> batch: multi-WFS step over dicts of device buffers."},
    {"batchTune", (PyCFunction)(void (*)(void))py_batchTune, METH_VARARGS | METH_KEYWORDS, "batchTune(batch, stream=0, reps=5): autotune the persistent launches."},
    {"batchRun", (PyCFunction)(void (*)(void))py_batchRun, METH_VARARGS | METH_KEYWORDS, "batchRun(batch, stream=0): one step (one graph launch)."},
    {"batchInfo", py_batchInfo, METH_VARARGS, "batchInfo(batch) -> (list of launch dicts, WFS outside them)."},
    {"batchSetConfig", (PyCFunction)(void (*)(void))py_batchSetConfig, METH_VARARGS | METH_KEYWORDS, "batchSetConfig(batch, i, waves_per_wg, workgroups, chunk=0)."},
    {"batchFree", py_batchFree, METH_VARARGS, "batchFree(batch)."},
    {"reconCreate", py_reconCreate, METH_VARARGS,
     "reconCreate(ncents, nacts, layout, reconmx_ptr, ld, index) -> handle of the device reconstructor step."},
    {"reconSetGains", py_reconSetGains, METH_VARARGS, "reconSetGains(r, gains, decay): floats or float32 arrays of nacts."},
    {"reconRun", (PyCFunction)(void (*)(void))py_reconRun, METH_VARARGS | METH_KEYWORDS,
     "reconRun(r, slopes_ptr, slope_stride, nframes, acts_ptr, acts_stride, stream=0): acts = decay*acts - gains*(data.reconmx)."},
    {"reconFree", py_reconFree, METH_VARARGS, "reconFree(r)."},
    {"atmosWorkDoubles", py_atmosWorkDoubles, METH_VARARGS, "atmosWorkDoubles(npup) -> doubles of device work space."},
    {"atmosPupilPhase", (PyCFunction)(void (*)(void))py_atmosPupilPhase, METH_VARARGS | METH_KEYWORDS,
     "atmosPupilPhase(layers, npup, intrinsic_ptr, pupfn_ptr, pupil_area, remove_piston, out_scale, out64_ptr, "
     "out32_ptr, work_ptr, stream=0): createPupilPhs on the device."},
    {"atmosLayerUpdate", (PyCFunction)(void (*)(void))py_atmosLayerUpdate, METH_VARARGS | METH_KEYWORDS,
     "atmosLayerUpdate(screen_ptr, rows, cols, oy, ox, cols_ptr, ncols, rows_ptr, nrows, col_add, row_add, stream=0) "
     "-> new (oy, ox): makeLayer on a ring-buffered screen."},
    {"dmSplineOperator", py_dmSplineOperator, METH_VARARGS,
     "dmSplineOperator(npup, nact, actoffset=0) -> (npup, nact) float64 natural-spline operator."},
    {"dmSurface", (PyCFunction)(void (*)(void))py_dmSurface, METH_VARARGS | METH_KEYWORDS,
     "dmSurface(act_ptr, nact, coupling, op_ptr, opT_ptr, npup, surface_ptr, work_ptr, stream=0): MirrorSurface.fit."},
    {"dmLineOfSight", (PyCFunction)(void (*)(void))py_dmLineOfSight, METH_VARARGS | METH_KEYWORDS,
     "dmLineOfSight(surface_ptr, dmpup, yoff, xoff, npup, factor, adjust, base64_ptr, base32_ptr, pupfn_ptr, "
     "pupil_sum, remove_piston, out32_ptr, work_ptr, stream=0): selectSubPupil on the device."},
    {nullptr, nullptr, 0, nullptr}};

static struct PyModuleDef centModule = {PyModuleDef_HEAD_INIT, "cent",
                                        "gfx950 Shack-Hartmann image formation and centroiding (drop-in for cmod.cent)",
                                        -1, centMethods};

PyMODINIT_FUNC PyInit_cent(void) {
  import_array();
  if (dasp_cent_abi_version() != DASP_CENT_ABI_VERSION) {
    PyErr_Format(PyExc_ImportError, "cent: engine objects have ABI version %d, this module expects %d (rebuild)",
                 dasp_cent_abi_version(), DASP_CENT_ABI_VERSION);
    return nullptr;
  }
  PyObject *m = PyModule_Create(&centModule);
  if (!m) return nullptr;
  CentError = PyErr_NewException("cent.error", nullptr, nullptr);
  Kept = PyDict_New();
  if (!CentError || !Kept) return nullptr;
  Py_INCREF(CentError);
  PyModule_AddObject(m, "error", CentError);
  Py_INCREF(Kept);
  PyModule_AddObject(m, "_kept", Kept);
  PyModule_AddIntConstant(m, "FORM_FFT32", DASP_CENT_FORM_FFT32);
  PyModule_AddIntConstant(m, "FORM_DFT_BF16", DASP_CENT_FORM_DFT_BF16);
  PyModule_AddIntConstant(m, "FORM_DFT_FP16", DASP_CENT_FORM_DFT_FP16);
  PyModule_AddIntConstant(m, "FORM_FFT32_GENERAL", DASP_CENT_FORM_FFT32_GENERAL);
  PyModule_AddIntConstant(m, "FORM_DFT_FP16_1", DASP_CENT_FORM_DFT_FP16_1);
  PyModule_AddIntConstant(m, "ABI_VERSION", DASP_CENT_ABI_VERSION);
  PyModule_AddObject(m, "NOALLOC_SCRATCH_BYTES", PyLong_FromLong(DASP_CENT_NOALLOC_SCRATCH_BYTES));
  return m;
}
