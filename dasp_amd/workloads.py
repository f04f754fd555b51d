"""The Shack-Hartmann workloads this engine is measured and checked on.

Each workload is a WFS geometry (telescope, pupil, sampling) plus the options of the centroid
engine, with seeded synthetic von Karman phase (science/infScrn.py:244-345 restated in
dasp_amd.synth). Phase is in radians at 640 nm for an r0 quoted at 500 nm.
"""
from dataclasses import dataclass, field

import numpy as np

from dasp_amd import synth
from dasp_amd.util.tel import obstructed_pupil


@dataclass
class WFS:
    name: str
    nsubx: int
    n: int                   # phase pixels per subaperture
    nfft: int
    nimg: int
    D: float                 # telescope diameter (m)
    obs: float               # central obstruction (fraction of D)
    r0: float                # at 500 nm (m)
    L0: float
    clipsize: int = 0        # 0: = nfft
    ncen: int = 0            # 0: = nimg
    sig: float = 1.0
    readnoise: float = 0.0
    addPoisson: int = 0
    noiseFloor: float = 0.0
    threshType: int = 0
    relativeFloor: bool = False   # engine option: threshType 0 floor = noiseFloor * image maximum
    correlation: int = 0     # correlationCentroiding
    parabolicFit: int = 0
    spot: bool = False       # per-subaperture LGS spot kernels
    spot_size: int = 0
    nframes: int = 1
    seed: int = 10
    extra: dict = field(default_factory=dict)

    @property
    def npup(self):
        return self.nsubx * self.n

    @property
    def clip(self):
        return self.clipsize or self.nfft

    @property
    def cen(self):
        return self.ncen or self.nimg

    def pupil(self):
        return obstructed_pupil(self.nsubx, self.n, self.obs, 0.5)

    def phase(self, frame=0):
        """(npup, npup) float32 phase of one frame (numpy, reproducible)."""
        return synth.phase_screen(self.npup, self.D, self.r0, self.L0, self.seed + 1000 * frame,
                                  wavelength_scale=500. / 640.)

    def spots(self):
        if not self.spot:
            return None
        P = self.spot_size or self.clip
        return synth.lgs_spots(self.nsubx, P, self.n, self.nfft, self.nimg, self.clip,
                               telDiam=self.D, pup=self.pupil().fn)


# The benchmark workloads: SCAO 8x8, 40x40 with noise, 80x80 nfft32, the multi-WFS set
# (LGS 80x80, NGS 8x8, truth 40x40) and the 80x80 Nyquist WFS with correlation.
SCAO8 = WFS("scao8x8", 8, 8, 16, 8, 4.2, 0.25, 0.15, 30.0)
NOISE40 = WFS("wfs40x40_noise", 40, 16, 32, 16, 8.0, 0.14, 0.12, 25.0, sig=1000.0, readnoise=0.5,
         addPoisson=1, noiseFloor=0.1, threshType=0, relativeFloor=True, nframes=64)
WFS80 = WFS("wfs80x80_nfft32", 80, 16, 32, 16, 39.0, 0.28, 0.157, 25.0, nframes=64)
LGS80 = WFS("lgs80x80", 80, 16, 32, 16, 39.0, 0.28, 0.157, 25.0, sig=1000.0, readnoise=0.5,
             addPoisson=1, spot=True)
NGS8 = WFS("ngs8x8", 8, 32, 64, 8, 39.0, 0.28, 0.157, 25.0, sig=1000.0, readnoise=0.5, addPoisson=1)
TRUTH40 = WFS("truth40x40", 40, 16, 32, 16, 39.0, 0.28, 0.157, 25.0, sig=1000.0, readnoise=0.5, addPoisson=1)
CORR80 = WFS("wfs80x80_nyquist_corr", 80, 32, 64, 16, 39.0, 0.28, 0.157, 25.0, correlation=1, parabolicFit=5,
         nframes=256)


def reference_images(w, pupil=None):
    """Flat-phase noiseless images of every subaperture normalised to a maximum of 1, the
    correlation reference util/centroid.py:1580-1626 builds (takeCorrImage)."""
    from dasp_amd.util import centcmod
    pupil = pupil or w.pupil()
    S = w.nsubx * w.nsubx
    cents = np.zeros((w.nsubx, w.nsubx, 2), np.float32)
    bimg = np.zeros((w.nsubx, w.nsubx, w.nimg, w.nimg), np.float32)
    phs = np.zeros((w.npup, w.npup), np.float32)
    cc = centcmod.centcmod(1, w.nsubx, w.cen, w.nfft, w.clip, w.nimg, w.n, 0., 0., 0, 0., float(w.sig), 0., 1, 1.,
                           1, 0, 1, phs, pupil, None, cents, bimg, 0.5, 0, None, 0, 0., None, None, 0, 1, 0, 1, 0,
                           None)
    cc.run(1)
    cc.free()
    ref = bimg.copy()
    mx = ref.max()
    return ref / mx if mx > 0 else ref
