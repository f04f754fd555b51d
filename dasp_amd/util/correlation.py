"""Correlation reference transform (util/correlation.py:20-74 of DASP).

transformPSF turns a reference spot image into the FFTW half-complex (R2HC x R2HC) form, with the
sign flips that conjugate it, that the engine's FFT correlation consumes (cmod/src/centmodule.c:
1403-1486).
"""
import numpy


def r2hc1d(a):
    """FFTW R2HC along the last axis: r_0..r_{n/2}, then i_{(n+1)/2-1}..i_1."""
    a = numpy.asarray(a, numpy.float64)
    n = a.shape[-1]
    res = numpy.fft.fft(a, axis=-1)
    b = numpy.empty_like(a)
    b[..., :n // 2 + 1] = res.real[..., :n // 2 + 1]
    if n // 2 + 1 < n:
        b[..., n // 2 + 1:] = res.imag[..., (n + 1) // 2 - 1:0:-1]
    return b


def r2hc(a):
    """Separable 2-D R2HC (rows, then columns)."""
    return numpy.swapaxes(r2hc1d(numpy.swapaxes(r2hc1d(a), -1, -2)), -1, -2)


def transformPSF(psf):
    """psf: (nimg,nimg) or (nsubx,nsubx,nimg,nimg) -> float32 half-complex conjugated spectra of
    fftshift(psf) (util/correlation.py:20-51)."""
    if psf is None:
        return None
    p = numpy.asarray(psf, numpy.float64)
    if p.ndim not in (2, 4):
        raise Exception("PSF shape wrong in util/correlation.py")
    ff = r2hc(numpy.fft.fftshift(p, axes=(-2, -1)))
    mm = ff.shape[-2] // 2 + 1
    nn = ff.shape[-1] // 2 + 1
    ff[..., :mm, nn:] *= -1
    ff[..., mm:, :nn] *= -1
    return ff.astype(numpy.float32)
