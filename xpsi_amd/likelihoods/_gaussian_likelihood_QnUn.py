"""Drop-in for xpsi/likelihoods/_gaussian_likelihood_QnUn.pyx.

gaussian_likelihood_QnUn (:14-44) keeps the reference signature and returns
(lnL, star): the Gaussian likelihood of every data row of counts [rows, bins],
bin by bin with the given errors, about the one model row star [bins] (the
normalised Stokes signal at the data phases).  The phase edges are taken and
not used, as in the reference.
"""
import ctypes as C

import numpy as np

from .._lib import lib, check
from .. import _device as D


def gaussian_likelihood_QnUn(phases, counts, errors, star):
    counts = np.ascontiguousarray(counts, dtype=np.float64)
    N_rows, N_b = counts.shape
    errors = np.ascontiguousarray(errors, dtype=np.float64)
    star = np.ascontiguousarray(star, dtype=np.float64)
    if errors.shape != (N_rows, N_b) or star.shape != (N_b,):
        raise TypeError("errors must match counts, and star must have one entry per bin")
    dcnt, derr, dst = D.to_dev(counts), D.to_dev(errors), D.to_dev(star)
    L = lib()
    wsb = L.xh_gaussian_workspace_bytes(N_rows, N_b)
    ws = D.empty((wsb + 7) // 8)
    lnl = C.c_double(0.0)
    rc = L.xh_gaussian_likelihood_QnUn(N_rows, N_b, D.ptr(dcnt), D.ptr(derr), D.ptr(dst), C.byref(lnl),
                                       D.ptr(ws), ws.numel() * 8, D.stream_ptr())
    check(rc, "xh_gaussian_likelihood_QnUn")
    return lnl.value, star
