"""Drop-in for xpsi/likelihoods/_gaussian_likelihood_given_background_IQU.pyx.

gaussian_likelihood_given_background (:14-91) keeps the reference signature
and returns (lnL, EXPEC): the Gaussian likelihood of the counts, bin by bin
with the given errors, about the expected counts EXPEC (star + background
count rates, phase-binned, times the exposure; the phase-bin integrals follow
xpsi/tools/compute_expected_counts.pyx:66-315).  Components may lie on
different phase grids and take allow_negative per component.
"""
import ctypes as C

import numpy as np

from .._lib import lib, check
from .. import _device as D
from ._components import allow_flags, pack


def gaussian_likelihood_given_background(exposure_time, phases, counts, errors, components,
                                         component_phases, phase_shifts, background, allow_negative=False):
    counts = np.ascontiguousarray(counts, dtype=np.float64)
    N_ch, N_b = counts.shape
    errors = np.ascontiguousarray(errors, dtype=np.float64)
    background = np.ascontiguousarray(background, dtype=np.float64)
    if errors.shape != (N_ch, N_b) or background.shape != (N_ch, N_b):
        raise TypeError("the errors and the background must have one entry per channel and phase interval")
    flags = allow_flags(allow_negative, len(components))
    shared, n_comp, N_P, dc, dcp, dnp = pack(components, component_phases, N_ch)
    dph, dcnt, derr = D.to_dev(phases), D.to_dev(counts), D.to_dev(errors)
    dsh = D.to_dev(np.asarray(phase_shifts, dtype=np.float64))
    dbg = D.to_dev(background)
    EXPEC = D.empty((N_ch, N_b))
    L = lib()
    wsb = L.xh_gaussian_workspace_bytes(N_ch, N_b)
    ws = D.empty((wsb + 7) // 8)
    lnl = C.c_double(0.0)
    head = (float(exposure_time), N_b, D.ptr(dph), N_ch, D.ptr(dcnt), D.ptr(derr), n_comp, D.ptr(dc))
    tail = (C.byref(lnl), D.ptr(EXPEC), D.ptr(ws), ws.numel() * 8, D.stream_ptr())
    if shared and len(set(flags)) == 1:
        rc = L.xh_gaussian_likelihood_given_background(*head, N_P, D.ptr(dcp), D.ptr(dsh), D.ptr(dbg),
                                                       int(flags[0]), *tail)
        check(rc, "xh_gaussian_likelihood_given_background")
    else:
        if shared:
            dnp = D.to_dev(np.full(n_comp, N_P, dtype=np.int32), np.int32)
            dcp = D.to_dev(np.tile(dcp.cpu().numpy(), n_comp))
        af = C.cast((C.c_int * n_comp)(*[int(f) for f in flags]), C.c_void_p)
        rc = L.xh_gaussian_likelihood_given_background_ex(*head, D.ptr(dnp), D.ptr(dcp), D.ptr(dsh), D.ptr(dbg),
                                                          af, *tail)
        check(rc, "xh_gaussian_likelihood_given_background_ex")
    return lnl.value, EXPEC.cpu().numpy()
