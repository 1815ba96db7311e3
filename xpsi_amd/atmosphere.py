"""Atmosphere tables on disk.

load_NSX_table follows xpsi/Photosphere.py:167-218: a text table whose rows
run over log T (outermost), log g, log10(E/kT) and mu (innermost, in
descending order), with columns chosen by Tcol, gcol, mucol, Ecol and
spe_Icol (log10 of I_nu / T^3).  It returns the axes (mu ascending) and the
intensities 10^(log I) flattened in C order (log T, log g, mu, log10(E/kT)),
the layout the 4-D table atmosphere takes.
"""
import numpy as np


def load_NSX_table(path, Tcol=4, gcol=5, mucol=1, Ecol=0, spe_Icol=2):
    table = np.loadtxt(path, dtype=np.double)
    nT = len(np.unique(table[:, Tcol]))
    ng = len(np.unique(table[:, gcol]))
    nmu = len(np.unique(table[:, mucol]))
    nE = len(np.unique(table[:, Ecol]))
    if table.shape[0] != nT * ng * nE * nmu:
        raise ValueError("the table is not a full grid over its four axes")
    t = table.reshape(nT, ng, nE, nmu, table.shape[1])
    logT = t[:, 0, 0, 0, Tcol].copy()
    logg = t[0, :, 0, 0, gcol].copy()
    logE = t[0, 0, :, 0, Ecol].copy()
    mu = t[0, 0, 0, ::-1, mucol].copy()
    # rows (T, g, E, mu descending) -> (T, g, mu ascending, E)
    buf = 10.0 ** t[:, :, :, ::-1, spe_Icol].transpose(0, 1, 3, 2)
    return logT, logg, mu, logE, np.ascontiguousarray(buf).reshape(-1)


class HotAtmosphere:
    """The integrators' atmosphere extension point on the device
    (hot_wrapper.pyx:72-252: init_hot / eval_hot_I / eval_hot_norm /
    free_hot), batched: ``eval_I(E, mu, VEC)`` evaluates a sequence of
    queries (E keV, mu, VEC = (log10 T, log10 g) per query) as one reference
    thread would in that order after init_hot.  ext 1: blackbody; ext 2: the
    4-D table given as (logT, logg, mu, logE, buf)."""

    def __init__(self, ext, table=None):
        import ctypes as C
        from ._lib import lib, XHError
        self._keep = []
        if ext == 2:
            arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in table]
            self._keep = arrs
            axes = (C.c_void_p * 4)(*[a.ctypes.data for a in arrs[:4]])
            N = (C.c_int * 4)(*[len(a) for a in arrs[:4]])
            self._h = lib().xh_atm_init(2, axes, N, arrs[4].ctypes.data)
        else:
            self._h = lib().xh_atm_init(int(ext), None, None, None)
        if not self._h:
            raise XHError("xh_atm_init: " + lib().xh_last_error().decode())

    def norm(self):
        from ._lib import lib
        return lib().xh_atm_norm(self._h)

    def eval_I(self, E, mu, VEC):
        from . import _device as D
        from ._lib import lib, check
        E = np.ascontiguousarray(E, dtype=np.float64)
        dE, dmu = D.to_dev(E), D.to_dev(np.array(np.broadcast_to(mu, E.shape), dtype=np.float64))
        dV = D.to_dev(np.ascontiguousarray(VEC, dtype=np.float64).reshape(len(E), 2))
        out = D.empty(len(E))
        check(lib().xh_atm_eval_I(self._h, len(E), D.ptr(dE), D.ptr(dmu), D.ptr(dV), D.ptr(out), D.stream_ptr()),
              "xh_atm_eval_I")
        return out.cpu().numpy()

    def close(self):
        from ._lib import lib
        if getattr(self, "_h", None):
            lib().xh_atm_free(self._h)
            self._h = None
