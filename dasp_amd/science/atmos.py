"""The atmosphere's pupil phase on the device: util/atmos.py ``atmos.createPupilPhs`` (:678-833) for a
natural guide star and science/infAtmos.py ``makeLayer`` (:364-402), over layer screens resident in
HBM (include/dasp_atmos.h, dasp_amd/csrc/atmos.hip).

The per-step bookkeeping stays the reference's: each layer's static offset (positionDict), its
add direction (colAdd / rowAdd) and this step's sub-pixel positions (interpPosCol / interpPosRow
from infAtmos's newCols/newRows) give a whole-pixel window start and two fractions
(util/atmos.py:697-725), the fractions give linearshift's float weights
(cmod/src/interpmodule.c:1276-1279); the device sums the layers' bilinear windows in float64,
removes the piston inside the pupil, applies the pupil and the 500/lambda scale.

makeLayer shifts the whole screen by the rows and columns the screen generator added and writes
them in; on the device the screen is a ring buffer instead, so the shift is a move of its origin and
only the new rows and columns are written.
"""
import math

import numpy as np

from dasp_amd.cmod import cent


def _f32(x):
    return float(np.float32(x))


def linearshift_weights(xshift, yshift):
    """interp_linearshift's weights: the shifts as C floats, products in double, stored as floats."""
    x = float(np.float32(xshift))
    y = float(np.float32(yshift))
    return _f32((1. - x) * (1. - y)), _f32((1. - x) * y), _f32(x * (1. - y)), _f32(x * y)


def window_shift(pos_start, add, interp_pos):
    """util/atmos.py:697-720 for one axis: (whole-pixel shift, fraction)."""
    p = pos_start
    if add >= 0:
        if interp_pos > 0:
            p += 1 - interp_pos
    else:
        if interp_pos > 0:
            p += interp_pos
        else:
            p += 1
    whole = int(math.floor(p))
    return whole, p - whole


class DeviceAtmos:
    """Layer screens in HBM and createPupilPhs on the device for one natural-guide-star direction.

    layers: list in increasing height of dicts with key, screen (2-D float64, the layer's
    phaseScreens[key]), x0, y0 (positionDict[key][0], [1]), pos_col, pos_row (positionDict[key][4],
    [5]) and col_add, row_add (the layer's colAdd / rowAdd). pupfn: the pupil function (npup x npup)
    or None for control["fullPupil"]; pupil_area defaults to its sum (pupil.area)."""

    def __init__(self, npup, layers, pupfn=None, pupil_area=None, source_lam=500., device="cuda:0"):
        import torch
        self.torch = torch
        self.dev = torch.device(device)
        self.npup = int(npup)
        self.layers = []
        for L in layers:
            scr = torch.as_tensor(np.ascontiguousarray(L["screen"], np.float64)).to(self.dev)
            self.layers.append(dict(key=L["key"], screen=scr, rows=scr.shape[0], cols=scr.shape[1], oy=0, ox=0,
                                    x0=int(L["x0"]), y0=int(L["y0"]), pos_col=float(L["pos_col"]),
                                    pos_row=float(L["pos_row"]), col_add=int(L["col_add"]),
                                    row_add=int(L["row_add"])))
        self.by_key = {L["key"]: L for L in self.layers}
        if pupfn is None:
            self.pupfn = None
            self.area = float(self.npup * self.npup)
        else:
            p = np.ascontiguousarray(pupfn, np.float32)
            self.pupfn = torch.from_numpy(p).to(self.dev)
            self.area = float(pupil_area) if pupil_area is not None else float(p.astype(np.float64).sum())
        self.source_lam = float(source_lam)
        self.work = torch.empty(cent.atmosWorkDoubles(self.npup), dtype=torch.float64, device=self.dev)

    def screen(self, key):
        """The layer's logical screen (phaseScreens[key]) as a device tensor (a rolled copy)."""
        L = self.by_key[key]
        return self.torch.roll(L["screen"], shifts=(-L["oy"], -L["ox"]), dims=(0, 1))

    def make_layer(self, key, new_cols, new_rows, stream=None):
        """makeLayer: new_cols (ncols, rows) as colInput[maxColAdd - ncols:], new_rows (nrows, cols) as
        rowInput[...]; float64 device tensors (or arrays, copied to the device)."""
        torch = self.torch
        L = self.by_key[key]
        nc = torch.as_tensor(np.ascontiguousarray(new_cols, np.float64) if isinstance(new_cols, np.ndarray) else new_cols,
                             dtype=torch.float64).to(self.dev).contiguous()
        nr = torch.as_tensor(np.ascontiguousarray(new_rows, np.float64) if isinstance(new_rows, np.ndarray) else new_rows,
                             dtype=torch.float64).to(self.dev).contiguous()
        st = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
        L["oy"], L["ox"] = cent.atmosLayerUpdate(L["screen"].data_ptr(), L["rows"], L["cols"], L["oy"], L["ox"],
                                                 nc.data_ptr(), nc.shape[0], nr.data_ptr(), nr.shape[0],
                                                 L["col_add"], L["row_add"], stream=st)
        L["_keep"] = (nc, nr)   # the copies stay alive until the update has run

    def create_pupil_phs(self, interp_col, interp_row, out64=None, out32=None, remove_piston=True,
                         intrinsic=None, stream=None):
        """createPupilPhs: interp_col / interp_row map each layer key to this step's
        interpPosColDict / interpPosRowDict value. out64 (float64) and / or out32 (float32) are
        (npup, npup) device tensors written with the result; returns out64 if given, else out32."""
        torch = self.torch
        if out64 is None and out32 is None:
            out64 = torch.empty((self.npup, self.npup), dtype=torch.float64, device=self.dev)
        desc = []
        for L in self.layers:
            sy, fc = window_shift(L["pos_col"], L["col_add"], interp_col[L["key"]])
            sx, fr = window_shift(L["pos_row"], L["row_add"], interp_row[L["key"]])
            s1, s2, s3, s4 = linearshift_weights(fc, fr)
            desc.append((L["screen"].data_ptr(), L["rows"], L["cols"], L["oy"], L["ox"], L["y0"] + sx, L["x0"] + sy,
                         s1, s2, s3, s4))
        st = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
        scale = 500. / self.source_lam if self.source_lam != 500. else 1.0
        cent.atmosPupilPhase(desc, self.npup, intrinsic.data_ptr() if intrinsic is not None else 0,
                             self.pupfn.data_ptr() if self.pupfn is not None else 0, self.area, bool(remove_piston),
                             scale, out64.data_ptr() if out64 is not None else 0,
                             out32.data_ptr() if out32 is not None else 0, self.work.data_ptr(), stream=st)
        return out64 if out64 is not None else out32
