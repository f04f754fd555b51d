"""The dense-matrix step of science/tomoRecon.py on the device, fed by slopes already in HBM.

science/tomoRecon.py builds the reconstructor's slope vector from the WFS it reconstructs from: the
x slopes of every active sub-aperture of every WFS, then all the y slopes (centIndex, :131-150;
getInput, :682-700), forms tmp = -data . reconmx (:1140-1146) and integrates
outputData = decay * outputData + gains * tmp[:nacts] (:1090-1107, :1174). ``DeviceRecon`` does the
same on the GPU (dasp_recon_* in include/dasp_cent.h): the slopes of several WFS sit in one device
buffer (``slope_buffer`` gives each WFS a view of it to pass to the engine as its slope output), the
gather, the product and the integrator run as two kernels, and nothing leaves HBM.
"""
import numpy

from dasp_amd.cmod import cent


def cent_index(subflags, offsets):
    """The reconstructor's slope order over WFS with the given sub-aperture flags (nsubx, nsubx),
    whose slopes (nsubx, nsubx, 2) start at the given float offsets of one buffer: first every x
    slope of every WFS, then every y slope (science/tomoRecon.py:143-150)."""
    xs, ys = [], []
    for flag, off in zip(subflags, offsets):
        ind = numpy.nonzero(numpy.asarray(flag).ravel())[0]
        xs.append(off + 2 * ind)
        ys.append(off + 2 * ind + 1)
    return numpy.ascontiguousarray(numpy.concatenate(xs + ys).astype(numpy.int32))


class DeviceRecon:
    """reconmx: a float32 device tensor, (ncents, nmodes) as tomoRecon's dot(data, reconmx)
    (layout 0) or (nmodes, ncents) (layout 1); only the first nacts modes are formed.
    subflags: per WFS (nsubx, nsubx) flags. gains / decay: floats or float32 arrays of nacts."""

    def __init__(self, subflags, reconmx, nacts=None, gains=1.0, decay=1.0, layout=0, frames=1, device="cuda:0"):
        import torch
        self.torch = torch
        self.dev = torch.device(device)
        self.nsubx = [int(numpy.asarray(f).shape[0]) for f in subflags]
        sizes = [2 * n * n for n in self.nsubx]
        self.offsets = numpy.concatenate([[0], numpy.cumsum(sizes)[:-1]]).astype(numpy.int64)
        self.slope_len = int(sum(sizes))
        self.index = cent_index(subflags, self.offsets)
        self.ncents = int(self.index.size)
        if reconmx.dtype != torch.float32 or not reconmx.is_contiguous() or reconmx.device != self.dev:
            raise ValueError("reconmx must be a contiguous float32 tensor on %s" % self.dev)
        rows, cols = reconmx.shape
        if layout == 0 and rows != self.ncents or layout == 1 and cols != self.ncents:
            raise ValueError("reconmx %s does not match %d slopes for layout %d" % (tuple(reconmx.shape), self.ncents, layout))
        self.nacts = int(nacts if nacts is not None else (cols if layout == 0 else rows))
        self.mx = reconmx
        self.layout = layout
        self.frames = int(frames)
        self.slopes = torch.zeros((self.frames, self.slope_len), dtype=torch.float32, device=self.dev)
        self.acts = torch.zeros((self.frames, self.nacts), dtype=torch.float32, device=self.dev)
        self.handle = cent.reconCreate(self.ncents, self.nacts, layout, reconmx.data_ptr(), cols, self.index)
        self.set_gains(gains, decay)

    def slope_buffer(self, k):
        """WFS k's slopes as a (frames, nsubx, nsubx, 2) view of the shared slope buffer."""
        n = self.nsubx[k]
        o = int(self.offsets[k])
        return self.slopes[:, o:o + 2 * n * n].view(self.frames, n, n, 2)

    def set_gains(self, gains, decay):
        g = numpy.ascontiguousarray(gains, numpy.float32) if isinstance(gains, numpy.ndarray) else float(gains)
        d = numpy.ascontiguousarray(decay, numpy.float32) if isinstance(decay, numpy.ndarray) else float(decay)
        cent.reconSetGains(self.handle, g, d)

    def run(self, nframes=None, stream=None):
        """acts[f] = decay * acts[f] - gains * (data[f] . reconmx) for every frame."""
        st = stream if stream is not None else self.torch.cuda.current_stream(self.dev).cuda_stream
        cent.reconRun(self.handle, self.slopes.data_ptr(), self.slope_len, int(nframes or self.frames),
                      self.acts.data_ptr(), self.nacts, st)

    def free(self):
        if self.handle:
            cent.reconFree(self.handle)
            self.handle = None
