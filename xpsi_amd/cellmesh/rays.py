"""Drop-in for xpsi/cellmesh/rays.pyx: compute_rays (:249-390).

Same signature and return tuple; maxAlpha (a numpy array) is clamped in place
as the reference does.  The lag is returned uncalibrated (the reference's
HotRegion calibrates it afterwards, HotRegion.py:880-901).
"""
import ctypes as C

import numpy as np

from .._lib import lib, check
from .. import _device as D


def compute_rays(N_T, N_u, r_s, r_s_over_R, maxAlpha, N_R):
    u = np.ascontiguousarray(r_s_over_R, dtype=np.float64)[:N_u]
    if not (isinstance(maxAlpha, np.ndarray) and maxAlpha.dtype == np.float64 and maxAlpha.flags.c_contiguous):
        raise TypeError("maxAlpha must be a C-contiguous float64 array (it is updated in place)")
    du, dmax = D.to_dev(u), D.to_dev(maxAlpha[:N_u])
    defl, cosa, lag = D.empty((N_u, N_R)), D.empty((N_u, N_R)), D.empty((N_u, N_R))
    maxD = D.empty(N_u)
    L = lib()
    wsb = L.xh_compute_rays_workspace_bytes(int(N_u), int(N_R))
    ws = D.empty((wsb + 7) // 8)
    term = C.c_int(0)
    rc = L.xh_compute_rays(int(N_u), float(r_s), D.ptr(du), D.ptr(dmax), int(N_R), D.ptr(defl), D.ptr(cosa),
                           D.ptr(lag), D.ptr(maxD), D.ptr(ws), ws.numel() * 8, C.byref(term), D.stream_ptr())
    check(rc, "xh_compute_rays")
    maxAlpha[:N_u] = dmax.cpu().numpy()
    return (term.value, defl.cpu().numpy(), cosa.cpu().numpy(), lag.cpu().numpy(), maxD.cpu().numpy())
