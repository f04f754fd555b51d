"""Multi-WFS step on one GPU: several centcmod instances, each with nframes device-resident
phase frames, run as one captured step graph (persistent work-queue launches per fused shape).

The reference steps each wavefront sensor separately (science/wfscent.py:630-733 calls
util/centroid.py runCalc, :561-615, which calls one centcmod.run per WFS and frame); this class
keeps the phase, pre-drawn noise, slopes and (optionally) images of every WFS in HBM and runs
them all with dasp_cent_batch_run.
"""
import numpy

from dasp_amd.cmod import cent


class WFSBatch:
    """Device buffers and a step graph for a list of centcmod instances.

    Per WFS k: phs[k] (nframes, npup, npup) float32, cents[k] (nframes, nsubx, nsubx, 2), and,
    when its noise is on, zpois[k]/zread[k] (nframes, nsubaps, nimg, nimg) pre-drawn N(0,1).
    Write new phase into phs[k] (in place) between steps; the step graph reads the same addresses.
    """

    def __init__(self, ccs, nframes, images=False, device="cuda:0"):
        import torch
        self.torch = torch
        self.dev = torch.device(device)
        self.ccs = list(ccs)
        self.nframes = int(nframes)
        self.phs, self.cents, self.bimg, self.zpois, self.zread = [], [], [], [], []
        items = []
        for cc in self.ccs:
            nsubx, nimg = cc.nsubx, cc.nimg
            npup = nsubx * cc.phasesize
            S = nsubx * nsubx
            phs = torch.zeros((nframes, npup, npup), dtype=torch.float32, device=self.dev)
            ce = torch.zeros((nframes, nsubx, nsubx, 2), dtype=torch.float32, device=self.dev)
            bi = torch.zeros((nframes, S, nimg, nimg), dtype=torch.float32, device=self.dev) if images else None
            pois = cc.addPoisson == 1
            rd = float(cc.readnoise) != 0.0
            zp = torch.zeros((nframes, S, nimg, nimg), dtype=torch.float32, device=self.dev) if pois else None
            zr = torch.zeros((nframes, S, nimg, nimg), dtype=torch.float32, device=self.dev) if rd else None
            self.phs.append(phs)
            self.cents.append(ce)
            self.bimg.append(bi)
            self.zpois.append(zp)
            self.zread.append(zr)
            items.append(dict(handle=cc.centstruct, phs=phs.data_ptr(), phs_stride=npup * npup, nframes=nframes,
                              cents=ce.data_ptr(), bimg=bi.data_ptr() if bi is not None else 0,
                              zpois=zp.data_ptr() if zp is not None else 0,
                              zread=zr.data_ptr() if zr is not None else 0))
        torch.cuda.synchronize(self.dev)
        self.batch = cent.batchCreate(items)

    def draw_noise(self, seed, readout=0):
        """Fill the noise arrays with the engine's noise streams: WFS k, frame f draws streams
        (seed + k, 2*(readout+f)) and (seed + k, 2*(readout+f) + 1), generated on the device."""
        st = self.stream()
        for k, cc in enumerate(self.ccs):
            for f in range(self.nframes):
                for arr, off in ((self.zpois[k], 0), (self.zread[k], 1)):
                    if arr is None:
                        continue
                    cent.drawNormalsDevice(seed + k, 2 * (readout + f) + off, arr[f].numel(), arr[f].data_ptr(), st)

    def stream(self):
        return self.torch.cuda.current_stream(self.dev).cuda_stream

    @staticmethod
    def _key(L):
        return "np%d_F%d_nimg%d_spot%d_fast%d_groups%d" % (L["np"], L["fftsize"], L["nimg"], L["spot"], L["fast"],
                                                          L["groups"])

    def tune(self, reps=5, cache=None):
        """Autotune every persistent launch. With cache (a JSON file path), launches whose key
        (shape, trig mode, queue length) is in the file take the stored configuration instead,
        and newly tuned ones are added to it."""
        import json
        import os
        stored = {}
        if cache and os.path.exists(cache):
            with open(cache) as f:
                stored = json.load(f)
        launches, _ = self.info()
        if stored and all(self._key(L) in stored for L in launches):
            for i, L in enumerate(launches):
                c = stored[self._key(L)]
                cent.batchSetConfig(self.batch, i, c["waves_per_wg"], c["workgroups"], c["chunk"])
            return self.info()
        cent.batchTune(self.batch, self.stream(), reps)
        launches, others = self.info()
        if cache:
            for L in launches:
                stored[self._key(L)] = dict(waves_per_wg=L["waves_per_wg"], workgroups=L["workgroups"],
                                            chunk=L["chunk"], tuned_ms=L["tuned_ms"])
            with open(cache, "w") as f:
                json.dump(stored, f, indent=1)
        return launches, others

    def run(self):
        cent.batchRun(self.batch, self.stream())

    def info(self):
        return cent.batchInfo(self.batch)

    def free(self):
        if self.batch:
            cent.batchFree(self.batch)
            self.batch = None

    def slopes(self, k):
        return self.cents[k].cpu().numpy()

    def subap_images_per_step(self):
        return self.nframes * sum(int(numpy.sum(cc.subflag == 1)) for cc in self.ccs)
