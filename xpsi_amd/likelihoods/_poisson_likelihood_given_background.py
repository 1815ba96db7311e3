"""Drop-in for xpsi/likelihoods/_poisson_likelihood_given_background.pyx.

poisson_likelihood_given_background (:14-113) keeps the reference signature
and returns (lnL, EXPEC): the Poisson likelihood of the counts given the
star's signal plus a known background, with EXPEC the expected counts
(star + background count rates, phase-binned, times the exposure).  The
phase-bin integrals follow xpsi/tools/compute_expected_counts.pyx:66-315;
components may lie on different phase grids and take allow_negative per
component.
"""
import ctypes as C

import numpy as np

from .._lib import lib, check
from .. import _device as D
from ._components import allow_flags, pack


def poisson_likelihood_given_background(exposure_time, phases, counts, components,
                                        component_phases, phase_shifts, background,
                                        neg_sum_ln_data_factorial=None, allow_negative=False):
    counts = np.ascontiguousarray(counts, dtype=np.float64)
    N_ch, N_b = counts.shape
    background = np.ascontiguousarray(background, dtype=np.float64)
    if background.shape != (N_ch, N_b):
        raise TypeError("the background must have one entry per channel and phase interval")
    flags = allow_flags(allow_negative, len(components))
    shared, n_comp, N_P, dc, dcp, dnp = pack(components, component_phases, N_ch)
    dph, dcnt = D.to_dev(phases), D.to_dev(counts)
    dsh = D.to_dev(np.asarray(phase_shifts, dtype=np.float64))
    dbg = D.to_dev(background)
    dpre = None if neg_sum_ln_data_factorial is None else D.to_dev(neg_sum_ln_data_factorial)
    EXPEC = D.empty((N_ch, N_b))
    L = lib()
    wsb = L.xh_poisson_workspace_bytes(N_ch)
    ws = D.empty((wsb + 7) // 8)
    lnl = C.c_double(0.0)
    head = (float(exposure_time), N_b, D.ptr(dph), N_ch, D.ptr(dcnt), n_comp, D.ptr(dc))
    tail = (C.byref(lnl), D.ptr(EXPEC), D.ptr(ws), ws.numel() * 8, D.stream_ptr())
    if shared and len(set(flags)) == 1:
        rc = L.xh_poisson_likelihood_given_background(*head, N_P, D.ptr(dcp), D.ptr(dsh), D.ptr(dbg), D.ptr(dpre),
                                                      int(flags[0]), *tail)
        check(rc, "xh_poisson_likelihood_given_background")
    else:
        if shared:
            dnp = D.to_dev(np.full(n_comp, N_P, dtype=np.int32), np.int32)
            dcp = D.to_dev(np.tile(dcp.cpu().numpy(), n_comp))
        af = C.cast((C.c_int * n_comp)(*[int(f) for f in flags]), C.c_void_p)
        rc = L.xh_poisson_likelihood_given_background_ex(*head, D.ptr(dnp), D.ptr(dcp), D.ptr(dsh), D.ptr(dbg),
                                                         D.ptr(dpre), af, *tail)
        check(rc, "xh_poisson_likelihood_given_background_ex")
    return lnl.value, EXPEC.cpu().numpy()
