"""Count data read from OGIP files: Data.from_evt (an event list binned in
channel and phase) and Data.from_pha (a spectrum), following
xpsi/Data.py:398-533 (same arguments, same binning and fields)."""
import numpy as np

from .ogip import hdu, read_fits


class Data:
    def __init__(self, counts, channels, phases, first, last, exposure_time):
        self.counts = np.ascontiguousarray(counts, dtype=np.float64)
        self.channels = np.asarray(channels)
        self.phases = np.asarray(phases, dtype=np.float64)
        self.first, self.last = int(first), int(last)
        self.exposure_time = float(exposure_time)

    @classmethod
    def from_evt(cls, path, n_phases=32, channels=None, phase_column="PULSE_PHASE", channel_column="PI",
                 phase_shift=0.0):
        ev = hdu(read_fits(path), "EVENTS")
        exposure = ev["header"]["EXPOSURE"]
        ch = np.asarray(ev["columns"][channel_column])
        ph = (np.asarray(ev["columns"][phase_column], dtype=np.float64) + phase_shift) % 1.0
        if channels is None:
            lo, hi = int(np.min(ch)), int(np.max(ch))
            channels = np.arange(lo, hi + 1)
        else:
            lo, hi = channels[0], channels[-1]
        ch_borders = np.arange(lo, hi + 2) - 0.5
        ph_borders = np.linspace(0.0, 1.0, n_phases + 1)
        mask = (ch >= lo) & (ch <= hi)
        counts, _, _ = np.histogram2d(ch[mask], ph[mask], bins=[ch_borders, ph_borders])
        return cls(counts.astype(np.double), channels=channels, phases=ph_borders, first=0, last=hi - lo,
                   exposure_time=exposure)

    @classmethod
    def from_pha(cls, path, channels=None):
        sp = hdu(read_fits(path), "SPECTRUM")
        h = sp["header"]
        exposure = float(h["EXPOSURE"])
        chd = np.asarray(sp["columns"]["CHANNEL"])
        if "COUNTS" in sp["columns"]:
            cnt = np.asarray(sp["columns"]["COUNTS"])
        else:
            cnt = np.asarray(sp["columns"]["RATE"])
            exposure = 1.0
        if channels is None:
            lo, hi = int(np.min(chd)), int(np.max(chd))
            channels = np.arange(lo, hi + 1)
        else:
            lo, hi = channels[0], channels[-1]
        cmap = dict(zip(chd.tolist(), cnt.tolist()))
        if not all(c in cmap for c in np.asarray(channels).tolist()):
            raise ValueError("Not all channels exist in channel_data.")
        counts = np.array([[float(cmap[c]) for c in np.asarray(channels).tolist()]], dtype=np.double).T
        d = cls(counts, channels=channels, phases=np.array([0.0, 1.0]), first=0, last=hi - lo,
                exposure_time=exposure)
        d.backscal = h.get("BACKSCAL")
        d.ancrfile = h.get("ANCRFILE")
        d.respfile = h.get("RESPFILE")
        if h.get("HDUCLAS2") == "TOTAL" and "BACKFILE" in h:
            d.backfile = h["BACKFILE"]
        return d
