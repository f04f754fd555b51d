"""Telescope pupil geometry used to derive the per-subaperture inputs of the centroid engine.

Restates the parts of DASP's util/tel.py that the centroiding path consumes: the circular,
centrally obstructed pupil (util/tel.py:37-53, :155-158) and the per-subaperture helpers
``perSubap`` (:597-613), ``getSubapFlag`` (:564-576) and ``getSubarea`` (:578-587).
Spiders, hexagonal segments and pupil maps are outside this package's scope.
"""
import numpy as np


def makeCircularGrid(nxpup, nypup=None, natype=np.float64, dosqrt=1, xoff=0, yoff=0):
    """Distance of each pixel centre from the array centre (util/tel.py:37-53)."""
    if nypup is None:
        nypup = nxpup
    tabx = np.arange(nxpup) - float(nxpup / 2.) + 0.5 - xoff
    taby = np.arange(nypup) - float(nypup / 2.) + 0.5 - yoff
    grid = tabx[np.newaxis, :] ** 2 + taby[:, np.newaxis] ** 2
    if dosqrt:
        grid = np.sqrt(grid)
    return grid.astype(natype)


class Pupil:
    """Annular pupil: fn = (r <= r1) & (r >= r2) on an npup x npup grid (util/tel.py:155-158).

    r1, r2 are the outer and inner radii in pixels; ``nsubx`` and ``minarea`` are the defaults of
    the per-subaperture helpers, as in the reference class."""

    def __init__(self, npup, r1=None, r2=0, nsubx=None, minarea=0.5):
        self.npup = int(npup)
        self.r1 = npup / 2. if r1 is None else r1
        self.r2 = r2
        self.nsubx = nsubx
        self.minarea = minarea
        grid = makeCircularGrid(self.npup)
        self.fn = np.logical_and(grid <= self.r1, grid >= self.r2)
        self.area = int(self.fn.sum())

    def perSubap(self, nsubx=None, vectorAlign=1):
        """(nsubx, nsubx, n, n[_v]) float32 per-subaperture pupil (util/tel.py:597-613)."""
        nsubx = self.nsubx if nsubx is None else nsubx
        if nsubx is None:
            raise ValueError("Number of x subaps must be defined")
        n = self.npup // nsubx
        nv = (n + 3) & ~3 if vectorAlign else n
        a = np.zeros((nsubx, nsubx, n, nv), np.float32)
        a[:, :, :, :n] = self.fn.reshape(nsubx, n, nsubx, n).transpose(0, 2, 1, 3)
        return a

    def getSubarea(self, nsubx):
        """Unvignetted pixel count per subaperture, float64 (util/tel.py:578-587)."""
        n = self.npup // nsubx
        return self.fn[:nsubx * n, :nsubx * n].reshape(nsubx, n, nsubx, n).sum(axis=(1, 3)).astype(np.float64)

    def getSubapFlag(self, nsubx, minarea=None):
        """1 where subarea >= minarea*n*n, int32 (util/tel.py:564-576)."""
        if minarea is None:
            minarea = self.minarea
        n = self.npup // nsubx
        return (self.getSubarea(nsubx) >= minarea * n * n).astype(np.int32)


def obstructed_pupil(nsubx, n, obstruction, minarea=0.5):
    """Pupil of diameter nsubx*n pixels with a central obstruction given as a fraction of the
    diameter, as DASP parameter files build it: Pupil(npup, npup/2, npup/2*obs)."""
    npup = nsubx * n
    return Pupil(npup, npup / 2., npup / 2. * obstruction, nsubx, minarea)
