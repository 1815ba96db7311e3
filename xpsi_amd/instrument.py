"""The instrument response: the fold of a registered signal through the
response matrix (xpsi/Instrument.py:163-197) on the device, and the matrix
read from OGIP files (Instrument.from_ogip_fits, xpsi/Instrument.py:388-537).

``Instrument(matrix, energy_edges, channels, channel_edges)`` keeps the
reference's fields; ``instrument(signal, irange, orange)`` returns
``matrix[orange[0]:orange[1], irange[0]:irange[1]] @ signal`` computed with
xh_fold; ``from_ogip_fits`` builds the response from an RMF (MATRIX and
EBOUNDS extensions, variable-length rows of F_CHAN/N_CHAN groups) times an
ARF (SPECRESP), or from an RSP alone, trims empty rows and columns as the
reference does, and records RMF, ARF, name and input-bin energies.
"""
import os

import numpy as np

from . import _device as D
from ._lib import check, lib
from .ogip import hdu, read_fits


class Instrument:
    def __init__(self, matrix, energy_edges, channels, channel_edges=None):
        self.matrix = np.ascontiguousarray(matrix, dtype=np.float64)
        self.energy_edges = np.ascontiguousarray(energy_edges, dtype=np.float64)
        self.channels = np.asarray(channels)
        self.channel_edges = None if channel_edges is None else np.asarray(channel_edges, dtype=np.float64)
        if self.matrix.shape != (len(self.channels), len(self.energy_edges) - 1):
            raise ValueError("the matrix must have one row per channel and one column per input bin")

    def __call__(self, signal, irange, orange):
        """matrix[orange, irange] @ signal on the device (signal [n_in, N_P])."""
        sig = np.ascontiguousarray(signal, dtype=np.float64)
        M = np.ascontiguousarray(self.matrix[orange[0]:orange[1], irange[0]:irange[1]])
        n_ch, n_in = M.shape
        if sig.shape[0] != n_in:
            raise ValueError("signal rows must match the input range")
        N_P = sig.shape[1]
        dM, dS = D.to_dev(M), D.to_dev(sig)
        out = D.empty((n_ch, N_P))
        check(lib().xh_fold(n_ch, n_in, N_P, D.ptr(dM), n_in, D.ptr(dS), D.ptr(out), D.stream_ptr()), "xh_fold")
        return out.cpu().numpy()

    @classmethod
    def from_ogip_fits(cls, RMF_path, ARF_path=None, min_channel=0, max_channel=-1, min_input=1,
                       max_input=-1, datafolder=None):
        if datafolder:
            ARF_path = os.path.join(datafolder, ARF_path) if ARF_path is not None else None
            RMF_path = os.path.join(datafolder, RMF_path)
        rmf = read_fits(RMF_path)
        mat, ebd = hdu(rmf, "MATRIX"), hdu(rmf, "EBOUNDS")
        hdr = mat["header"]
        instr = hdr.get("INSTRUME")
        detchans, numgrp = int(hdr["DETCHANS"]), int(hdr["NAXIS2"])
        fchan_col = [k for k in range(1, hdr["TFIELDS"] + 1) if hdr.get("TTYPE%d" % k) == "F_CHAN"][0]
        tlmin = int(hdr.get("TLMIN%d" % fchan_col, 1))
        tlmax = int(hdr.get("TLMAX%d" % fchan_col, detchans - 1 + tlmin))
        if ARF_path is not None:
            arf = hdu(read_fits(ARF_path), "SPECRESP")
            if arf["header"].get("INSTRUME") != instr:
                raise ValueError("the ARF and RMF belong to different instruments")
        if min_channel == 0:
            min_channel = tlmin
        if max_channel == -1:
            max_channel = detchans - 1
        if max_input == -1:
            max_input = numgrp
        if not (min_channel >= tlmin and max_channel <= tlmax):
            raise ValueError("channel range outside TLMIN..TLMAX")
        if not (min_input >= 0 and max_input <= numgrp):
            raise ValueError("input range outside the matrix")
        c = mat["columns"]
        channels = np.asarray(ebd["columns"]["CHANNEL"])
        inputs = np.arange(1, numgrp + 1)
        RMF = np.zeros((detchans, numgrp))
        for i in range(numgrp):
            n_grp = int(np.atleast_1d(c["N_GRP"][i])[0]) if np.ndim(c["N_GRP"][i]) else int(c["N_GRP"][i])
            if n_grp == 0:
                continue
            f = np.atleast_1d(c["F_CHAN"][i])
            n = np.atleast_1d(c["N_CHAN"][i])
            row = np.atleast_1d(c["MATRIX"][i])
            skip = 0
            for fc, nc in zip(f, n):
                fc, nc = int(fc), int(nc)
                if nc == 0:
                    continue
                RMF[fc - tlmin:fc + nc - tlmin, i] += row[skip:skip + nc]
                skip += nc
        ci = np.where((channels >= min_channel) & (channels <= max_channel))[0]
        ii = np.where((inputs >= min_input) & (inputs <= max_input))[0]
        channels, inputs = channels[ci], inputs[ii]
        RMF = RMF[ci][:, ii]
        if ARF_path is None:
            RSP = RMF
            area = RMF.sum(axis=0)
        else:
            area = np.asarray(arf["columns"]["SPECRESP"], dtype=np.float64)[ii]
            RSP = RMF * area
        empty_ch = np.all(RSP == 0, axis=1)
        empty_in = np.all(RSP == 0, axis=0)
        RSP = RSP[~empty_ch][:, ~empty_in]
        channels, inputs = channels[~empty_ch], inputs[~empty_in]
        elo, ehi = np.asarray(c["ENERG_LO"], np.float64), np.asarray(c["ENERG_HI"], np.float64)
        edges = np.append(elo[inputs - 1], ehi[inputs[-1] - 1])
        e_min, e_max = np.asarray(ebd["columns"]["E_MIN"], np.float64), np.asarray(ebd["columns"]["E_MAX"], np.float64)
        ch_edges = np.append(e_min[channels - tlmin], e_max[channels[-1] - tlmin])
        inst = cls(RSP, edges, channels, ch_edges)
        inst.RMF = RMF[~empty_ch][:, ~empty_in]
        inst.ARF = area[~empty_in]
        inst.name = instr
        inst.energies = ((elo + ehi) / 2.0)[ii][~empty_in]
        return inst
