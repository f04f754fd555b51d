"""Synthetic inputs of the Shack-Hartmann path: von Karman phase screens and LGS spot kernels.

The phase screen restates DASP's FFT screen generator science/infScrn.py:244-345 (makeInitialScreen,
called through makeScrnQuick :241-243). The elongated LGS spots restate util/elong.py:22-52 (Gauss)
and :418-486 (make), with the pixel scale of util/calcPxlScale.py:24-41.
"""
import math

import numpy as np


def _vk_filter(nfft, scrn_size, r0, L0):
    """Amplitude filter of the von Karman spectrum on an nfft^2 grid (science/infScrn.py:293-312)."""
    axe = np.arange(nfft) - nfft / 2.
    f = np.sqrt(axe[:, None] ** 2 + axe[None, :] ** 2)
    ff = np.fft.fftshift(f)
    fact = math.gamma(11. / 6.) ** 2 / (2 * math.pi ** (11. / 3.)) * ((24. / 5) * math.gamma(6. / 5)) ** (5. / 6)
    f0 = 1. / L0 * scrn_size
    filt = (ff ** 2 + f0 ** 2) ** (-11. / 12.)
    filt *= math.sqrt(fact * (scrn_size / r0) ** (5. / 3.)) * nfft
    return filt


def _screen_geometry(npup, D, L0):
    dpix = npup + 1                          # makeScrnQuick: dpix = npup+1, then [1:-1,1:-1]
    if L0 >= D:
        scrn_size = 2 * L0
    elif D >= 2 * L0:
        scrn_size = D
    else:
        scrn_size = 2 * L0
    nfft = int(np.around(dpix * 1. * scrn_size / D))
    nfft = max(nfft, dpix)
    return dpix, scrn_size, nfft


def phase_screen(npup, D, r0, L0, seed, wavelength_scale=1.0):
    """(npup, npup) float32 von Karman phase in radians (numpy, legacy MT19937 stream).

    r0 is at the wavelength the screen is quoted at; wavelength_scale multiplies the phase
    (e.g. 500/640 to view an r0@500nm screen at 640 nm)."""
    dpix, scrn_size, nfft = _screen_geometry(npup, D, L0)
    rn = np.random.RandomState(seed).randn(nfft, nfft)
    filt = _vk_filter(nfft, scrn_size, r0, L0)
    phi = np.fft.ifft2(np.fft.fft2(rn) * filt).real
    # phaseArray[:-1,:-1] = phi[:dpix,:dpix] (wind from the default direction), then crop
    full = np.zeros((dpix + 1, dpix + 1))
    full[:-1, :-1] = phi[:dpix, :dpix]
    return (full[1:-1, 1:-1] * wavelength_scale).astype(np.float32)


def phase_screens_torch(nframes, npup, D, r0, L0, seed, device, wavelength_scale=1.0):
    """(nframes, npup, npup) float32 von Karman phase generated on a torch device.

    Same spectrum and crop as phase_screen, with the white noise drawn by torch's generator
    (seeded) so that large batches are generated on the GPU in milliseconds."""
    import torch
    dpix, scrn_size, nfft = _screen_geometry(npup, D, L0)
    filt = torch.as_tensor(_vk_filter(nfft, scrn_size, r0, L0), dtype=torch.float32, device=device)
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    out = torch.empty((nframes, npup, npup), dtype=torch.float32, device=device)
    for k in range(nframes):
        rn = torch.randn((nfft, nfft), generator=g, device=device, dtype=torch.float32)
        phi = torch.fft.ifft2(torch.fft.fft2(rn) * filt).real
        out[k] = phi[1:npup + 1, 1:npup + 1] * wavelength_scale
    return out


def pxl_scale(lam, diam, n, nfft, binfactor):
    """Arcsec per pixel (util/calcPxlScale.py:24-41)."""
    if lam > 1:
        lam *= 1e-9
    return lam / diam * n / nfft * binfactor * 180 * 3600 / math.pi


def gauss_spot(n, fwhmx, fwhmy, th_deg, xoff=0., yoff=0.):
    """Rotated elliptical Gaussian on an n x n grid (util/elong.py:22-52)."""
    sig2x = fwhmx ** 2 / (8. * math.log(2.))
    sig2y = fwhmy ** 2 / (8. * math.log(2.))
    thr = th_deg * math.pi / 180.
    a = math.cos(thr) ** 2 / (2. * sig2x) + math.sin(thr) ** 2 / (2. * sig2y)
    b = math.sin(2. * thr) / (4. * sig2y) - math.sin(2. * thr) / (4. * sig2x)
    c = math.sin(thr) ** 2 / (2. * sig2x) + math.cos(thr) ** 2 / (2. * sig2y)
    i, j = np.mgrid[0:n, 0:n].astype(np.float64)
    x = (i - n // 2) + 0.5 - xoff
    y = (j - n // 2) + 0.5 - yoff
    return np.exp(-1. * (c * y * y + 2. * b * x * y + a * x * x))


def lgs_spots(nsubx, spotsize, wfs_n, wfs_nfft, wfs_nimg, clipsize=None, lam=640e-9, telDiam=39.,
              unelong_width=1.0, elong_min=0.5, elong_max=3.0, launch_dist=0., launch_theta=0.,
              pup=None):
    """(nsubx, nsubx, spotsize, spotsize) float32 elongated Gaussian spot kernels.

    Follows util/elong.py:418-486: the short axis is ``unelong_width`` arcsec; the long axis is
    sqrt(da^2 + w^2) rotated to the radial direction from the launch position. Instead of the
    sodium-layer geometry (beacon altitude/depth), da grows linearly with the subaperture's
    distance from the launch point so that the long axis spans elong_min..elong_max arcsec
    (FWHM) across the pupil. Each kernel is normalised to sum to the subaperture's lit fraction,
    as the reference scales by pup.sum()/spotsize^2."""
    if clipsize is None:
        clipsize = wfs_nfft
    binfactor = clipsize / wfs_nimg
    pixscale = pxl_scale(lam, float(telDiam) / nsubx, wfs_n, wfs_nfft, clipsize / wfs_nimg) / binfactor
    th0 = launch_theta - 90
    lx = launch_dist * math.cos(th0 / 180 * math.pi)
    ly = launch_dist * math.sin(th0 / 180 * math.pi)
    rmax = 0.0
    geo = []
    for isub in range(nsubx):
        for jsub in range(nsubx):
            xs = (float(isub) - nsubx / 2. + 0.5) * telDiam / nsubx + lx
            ys = (float(jsub) - nsubx / 2. + 0.5) * telDiam / nsubx + ly
            r = math.hypot(xs, ys)
            rmax = max(rmax, r)
            geo.append((isub, jsub, r, math.atan2(xs, ys) * 180. / math.pi))
    psfs = np.zeros((nsubx, nsubx, spotsize, spotsize), np.float32)
    for isub, jsub, r, phi in geo:
        e = elong_min + (elong_max - elong_min) * (r / rmax if rmax > 0 else 0.0)
        g = gauss_spot(spotsize, unelong_width / pixscale, e / pixscale, phi)
        frac = 1.0
        if pup is not None:
            n = pup.shape[0] // nsubx
            frac = pup[isub * n:(isub + 1) * n, jsub * n:(jsub + 1) * n].sum() / float(n * n)
        psfs[isub, jsub] = g * (frac / g.sum())
    return psfs
