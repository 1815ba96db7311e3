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
