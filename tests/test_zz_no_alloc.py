"""GPU: dasp_cent_run_device_noalloc against dasp_cent_run_device, both called through the C ABI.

The allocation-free entry point takes its scratch from one fixed device array instead of the
handle's buffers. For each path that uses scratch (correlation stage with no image buffers,
general path with its image-formation buffer, phase interpolation) two calls in a row, the second
over fewer frames and new phase, must give exactly what the allocating entry point gives; a call
whose scratch exceeds the array is refused before anything runs.
"""
import ctypes

import numpy as np
import pytest

from dasp_amd import workloads as W
from tests._cases import Case, say

pytestmark = pytest.mark.gpu


def _abi():
    from dasp_amd.cmod import cent
    lib = ctypes.CDLL(cent.__file__)   # the loaded extension: the same engine and handles
    vp = ctypes.c_void_p
    for name in ("dasp_cent_run_device", "dasp_cent_run_device_noalloc"):
        f = getattr(lib, name)
        f.argtypes = [vp, vp, ctypes.c_long, ctypes.c_int, vp, vp, vp, vp, vp, vp]
        f.restype = ctypes.c_int
    lib.dasp_cent_last_error.restype = ctypes.c_char_p
    return lib


def _run(lib, fn, c, phs, nframes, bimg=None, corr=None):
    import torch
    S = c.nsubx * c.nsubx
    ce = torch.full((nframes, S, 2), -9.0, dtype=torch.float32, device="cuda:0")
    stride = phs[0].numel()
    rc = getattr(lib, fn)(c.cc.centstruct, phs.data_ptr(), stride, nframes, ce.data_ptr(),
                          bimg.data_ptr() if bimg is not None else None,
                          corr.data_ptr() if corr is not None else None, None, None,
                          torch.cuda.current_stream().cuda_stream)
    assert rc == 0, lib.dasp_cent_last_error().decode()
    torch.cuda.synchronize()
    return ce.cpu().numpy()


def _phases(w, frames, npup, scale, m=None):
    import torch
    side = m or npup
    out = []
    for f in range(frames):
        p = W.WFS.phase(w, f) * scale
        if m is not None:
            p = np.ascontiguousarray(p[:m, :m])
        out.append(p)
    return torch.from_numpy(np.stack(out)).to("cuda:0").reshape(frames, side, side)


def _cases():
    w32 = W.WFS("na32", 8, 32, 64, 16, 8.0, 0.2, 0.15, 25.0)
    corr = dict(correlationCentroiding=1, corrPattern=W.reference_images(w32), parabolicFit=5)
    return {
        # correlation stage with no image or correlation-image buffer: both images in scratch
        "corr_stage_fft32": (w32, lambda: Case(8, 32, 64, 16, obs=0.2, form=10, **corr), None),
        "corr_stage_dft_fp16": (w32, lambda: Case(8, 32, 64, 16, obs=0.2, form=12, **corr), None),
        # fractional binning: the general path (image-formation buffer and images in scratch)
        "general_path": (W.WFS("nag", 6, 8, 16, 6, 4.2, 0.25, 0.2, 30.0),
                         lambda: Case(6, 8, 16, 6, clip=14, obs=0.25), None),
        # a 45 x 45 phase grid resampled onto 48 x 48 before the fused kernel
        "interpolation": (W.WFS("nai", 6, 8, 16, 8, 4.2, 0.25, 0.2, 30.0),
                          lambda: Case(6, 8, 16, 8, obs=0.25, phs=np.zeros((45, 45), np.float32)), 45),
    }


@pytest.mark.parametrize("name", ["corr_stage_fft32", "corr_stage_dft_fp16", "general_path", "interpolation"])
def test_noalloc_equals_run_device(name):
    lib = _abi()
    w, make, m = _cases()[name]
    c = make()
    npup = c.nsubx * c.n
    for calls, (nframes, scale) in enumerate(((3, 1.0), (2, -0.7))):   # the second call is smaller
        say("case no-alloc %s: call %d, %d frames" % (name, calls, nframes))
        phs = _phases(w, nframes, npup, scale, m)
        want = _run(lib, "dasp_cent_run_device", c, phs, nframes)
        got = _run(lib, "dasp_cent_run_device_noalloc", c, phs, nframes)
        act = c.cc.subflag.ravel() == 1
        assert np.abs(want[:, act]).max() > 0
        d = np.abs(got - want).max()
        say("  max|noalloc - run_device| = %.3g px" % d)
        np.testing.assert_array_equal(got, want)
    c.free()


def test_noalloc_refuses_oversized_scratch():
    """Images and correlation images of 129 frames of a 16x16 n32 WFS (0.5 MB per frame) exceed the
    64 MiB array: the call fails before launching anything, and a call that fits still works."""
    import torch
    from dasp_amd.cmod import cent
    lib = _abi()
    w = W.WFS("nabig", 16, 32, 64, 16, 8.0, 0.2, 0.15, 25.0)
    c = Case(16, 32, 64, 16, obs=0.2, form=10, correlationCentroiding=1, corrPattern=W.reference_images(w),
             parabolicFit=5)
    S, npup = 256, 512
    frames = 129
    say("case no-alloc oversized: %d frames" % frames)
    assert frames * S * (16 * 16 + 16 * 16) * 4 > cent.NOALLOC_SCRATCH_BYTES
    phs = torch.zeros((frames, npup, npup), dtype=torch.float32, device="cuda:0")
    ce = torch.zeros((frames, S, 2), dtype=torch.float32, device="cuda:0")
    rc = lib.dasp_cent_run_device_noalloc(c.cc.centstruct, phs.data_ptr(), npup * npup, frames, ce.data_ptr(),
                                          None, None, None, None, torch.cuda.current_stream().cuda_stream)
    msg = lib.dasp_cent_last_error().decode()
    say("  refused: rc=%d %s" % (rc, msg))
    assert rc != 0 and "scratch" in msg
    # with the caller's own image buffers nothing is needed from the array
    bi = torch.zeros((frames, S, 16, 16), dtype=torch.float32, device="cuda:0")
    ci = torch.zeros((frames, S, 16, 16), dtype=torch.float32, device="cuda:0")
    phs[:2].copy_(torch.from_numpy(np.stack([w.phase(0), w.phase(1)])))
    got = _run(lib, "dasp_cent_run_device_noalloc", c, phs[:2], 2, bi[:2], ci[:2])
    want = _run(lib, "dasp_cent_run_device", c, phs[:2], 2)
    np.testing.assert_array_equal(got, want)
    c.free()
