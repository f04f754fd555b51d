"""C-ABI checks that need no GPU: the C library and the cent extension load, the library exports
every symbol include/dasp_cent.h declares, the extension imports under the reference's name
cmod.cent, its host-only helpers agree with the oracle, and GPU entry points fail loudly here."""
import ctypes
import os
import re
import subprocess
import sys
import sysconfig

import numpy as np
import pytest

from oracle import centoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _lib():
    from dasp_amd.cmod import cent
    return cent, ctypes.CDLL(os.path.join(ROOT, "dasp_amd", "libdasp_cent.so"))


def test_exports_every_declared_symbol():
    """Both the C library and the extension module carry every entry point of every header under
    include/ (dasp_cent.h: engine, multi-WFS step, reconstructor; dasp_atmos.h: phase producer)."""
    import glob
    cent, lib = _lib()
    names = set()
    for h in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))):
        hdr = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)   # declarations, not comments
        names |= set(re.findall(r"\b(dasp_\w+)\s*\(", hdr))
    names = sorted(names)
    assert len(names) >= 25 and "dasp_atmos_pupil_phase" in names and "dasp_dm_surface" in names
    ext = ctypes.CDLL(cent.__file__)
    for nm in names:
        assert hasattr(lib, nm), nm
        assert hasattr(ext, nm), nm
    assert lib.dasp_cent_abi_version() == 2 == cent.abiVersion() == cent.ABI_VERSION


def test_dm_spline_operator_host_and_argument_checks():
    """The host-built natural-spline operator equals the oracle's splines of unit vectors; the
    device entry points refuse bad arguments before touching a GPU."""
    from dasp_amd.cmod import cent
    from oracle import atmosoracle as A
    for npup, nact in ((33, 9), (1280, 81), (98, 85)):
        op = cent.dmSplineOperator(npup, nact)
        x, x2 = A.spline_coords(npup, nact)
        I = np.eye(nact)
        ref = np.stack([A.natural_spline(x2, I[a], x) for a in range(nact)], axis=1)
        assert np.abs(op - ref).max() < 1e-12
    with pytest.raises(cent.error):
        cent.dmSplineOperator(64, 9, 0.5)        # pixels outside the actuator span (GSL refuses them)
    with pytest.raises(cent.error):
        cent.atmosPupilPhase([], 16, 0, 0, 256.0, True, 1.0, 0, 0, 0)   # no output
    with pytest.raises(cent.error):
        cent.atmosLayerUpdate(0, 10, 10, 0, 0, 0, 0, 0, 0, -1, -1)      # no screen
    assert cent.atmosWorkDoubles(1280) == 1280 * 1280 + 1280 * 5 + 1


def test_extension_imports_as_cmod_cent(tmp_path):
    """The built extension is the module DASP imports as cmod.cent (util/centcmod.py:16): dropped
    into a cmod package under its own file name it imports with no Python shim, has the six
    reference functions and cent.error, and runs a host-only function."""
    from dasp_amd.cmod import cent
    pkg = tmp_path / "cmod"
    pkg.mkdir()
    (pkg / "__init__.py").write_text("")
    os.symlink(cent.__file__, pkg / ("cent" + sysconfig.get_config_var("EXT_SUFFIX")))
    code = ("import numpy as np, cmod.cent as c\n"
            "assert all(hasattr(c, f) for f in ('initialise', 'run', 'runSlope', 'update', 'free', 'testcorrelation'))\n"
            "assert issubclass(c.error, Exception) and c.abiVersion() == 2\n"
            "img = np.zeros((8, 8), np.float32); img[3, 4] = 1\n"
            "corr = np.zeros((8, 8), np.float32); c.testcorrelation(img, corr)\n"
            "assert not corr.any()\n"
            "try:\n"
            "    c.run(12345)\n"
            "except c.error as e:\n"
            "    print('ok', e)\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=str(tmp_path), capture_output=True, text=True,
                       env=dict(os.environ, PYTHONPATH=str(tmp_path)))
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.startswith("ok invalid centstruct handle"), r.stdout


def test_noise_stream_is_deterministic_normal():
    cent, _ = _lib()
    a = cent.drawNormals(7, 3, 200001)
    b = cent.drawNormals(7, 3, 200001)
    c = cent.drawNormals(7, 4, 1000)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a[:1000], c)
    assert abs(a.mean()) < 0.01 and abs(a.std() - 1) < 0.01
    assert abs((a ** 3).mean()) < 0.03 and abs((a ** 4).mean() - 3) < 0.05


def test_noise_stream_known_answer():
    """The generator is PCG32 XSH-RR 64/32 (O'Neill 2014): its published demo sequence for
    initstate 42, initseq 54 starts 0xa15c02b7 0x7b47f409 0xba1d3330 0x83d2f293 0xbfa4784b
    0xcbed606e; consecutive outputs (u1, u2) = ((x + 0.5) / 2^32) give the Box-Muller pair
    r cos(2 pi u2), r sin(2 pi u2), r = sqrt(-2 ln u1)."""
    cent, _ = _lib()
    x = np.array([0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e], np.float64)
    u = (x + 0.5) / 2.0 ** 32
    r = np.sqrt(-2.0 * np.log(u[0::2]))
    th = 2.0 * np.pi * u[1::2]
    want = np.stack([r * np.cos(th), r * np.sin(th)], axis=1).ravel().astype(np.float32)
    got = cent.drawNormals(42, 54, 6)
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=0)
    np.testing.assert_array_equal(cent.drawNormals(42, 54, 5), got[:5])   # odd n: the last pair's cosine


@pytest.mark.parametrize("C", [4, 6, 7, 8, 16, 32])
def test_hc_to_spectrum_matches_oracle(C):
    cent, _ = _lib()
    rng = np.random.default_rng(C)
    hc = O.transform_psf(rng.random((C, C)))
    K = cent.hcToSpectrum(hc)
    K2 = O.hc_pattern_to_spectrum(hc.astype(np.float64))
    np.testing.assert_allclose(K, K2, atol=2e-6 * np.abs(K2).max())


@pytest.mark.parametrize("N", [8, 16])
def test_testcorrelation_matches_oracle(N):
    cent, _ = _lib()
    rng = np.random.default_rng(N)
    img = rng.random((N, N)).astype(np.float32)
    ref = rng.random((N, N))
    corr = O.transform_psf(ref)
    want = O.calc_correlation_hc(img.astype(np.float64), corr.astype(np.float64), N, N)
    cent.testcorrelation(img, corr)
    np.testing.assert_allclose(corr, want, rtol=1e-5, atol=1e-4)


def test_initialise_rejects_bad_arrays_before_touching_the_gpu():
    cent, _ = _lib()
    nsubx, n = 2, 8
    pupfn = np.ones((nsubx, nsubx, n, n), np.float32)
    cents = np.zeros((nsubx, nsubx, 2), np.float64)        # wrong dtype
    with pytest.raises(cent.error):
        cent.initialise(1, 4, 8, 16, 16, 8, n, 0., 0., 0, 0., 1.0, 0., 0, 1., 1, 0, 1,
                        np.zeros((16, 16), np.float32), pupfn, None, cents, np.ones((2, 2), np.int32),
                        np.zeros((2, 2, 8, 8), np.float32), np.ones((2, 2), np.float32), 0, None, 0, 0.,
                        None, None, 0, 0, 0, 1, 0, 0, 0., 0., None, None)


def test_no_silent_cpu_fallback_without_gpu():
    """With no usable gfx950 device the engine refuses to initialise; it never computes on the CPU."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    cent, _ = _lib()
    nsubx, n = 2, 8
    with pytest.raises(cent.error):
        cent.initialise(1, 4, 8, 16, 16, 8, n, 0., 0., 0, 0., 1.0, 0., 0, 1., 1, 0, 1,
                        np.zeros((16, 16), np.float32), np.ones((2, 2, n, n), np.float32), None,
                        np.zeros((2, 2, 2), np.float32), np.ones((2, 2), np.int32),
                        np.zeros((2, 2, 8, 8), np.float32), np.ones((2, 2), np.float32), 0, None, 0, 0.,
                        None, None, 0, 0, 0, 1, 0, 0, 0., 0., None, None)
