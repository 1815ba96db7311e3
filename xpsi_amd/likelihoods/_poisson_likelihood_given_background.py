"""Drop-in for xpsi/likelihoods/_poisson_likelihood_given_background.pyx.

poisson_likelihood_given_background (:14-113) keeps the reference signature
and returns (lnL, EXPEC): the Poisson likelihood of the counts given the
star's signal plus a known background, with EXPEC the expected counts
(star + background count rates, phase-binned, times the exposure).  The
phase-bin integrals follow xpsi/tools/compute_expected_counts.pyx:66-315.
Components must share one phase grid.
"""
import ctypes as C

import numpy as np

from .._lib import lib, check
from .. import _device as D


def poisson_likelihood_given_background(exposure_time, phases, counts, components,
                                        component_phases, phase_shifts, background,
                                        neg_sum_ln_data_factorial=None, allow_negative=False):
    counts = np.ascontiguousarray(counts, dtype=np.float64)
    N_ch, N_b = counts.shape
    background = np.ascontiguousarray(background, dtype=np.float64)
    if background.shape != (N_ch, N_b):
        raise TypeError("the background must have one entry per channel and phase interval")
    ph0 = np.asarray(component_phases[0], dtype=np.float64)
    for cp in component_phases[1:]:
        if not np.array_equal(np.asarray(cp), ph0):
            raise NotImplementedError("components must share one phase grid")
    if not isinstance(allow_negative, bool):
        if len(set(bool(a) for a in allow_negative)) != 1:
            raise NotImplementedError("per-component allow_negative is not supported")
        allow_negative = bool(allow_negative[0])
    comps = np.ascontiguousarray(np.stack([np.asarray(c, dtype=np.float64) for c in components]))
    n_comp, _, N_P = comps.shape
    dc, dph, dcnt = D.to_dev(comps), D.to_dev(phases), D.to_dev(counts)
    dcp, dsh = D.to_dev(ph0), D.to_dev(np.asarray(phase_shifts, dtype=np.float64))
    dbg = D.to_dev(background)
    dpre = None if neg_sum_ln_data_factorial is None else D.to_dev(neg_sum_ln_data_factorial)
    EXPEC = D.empty((N_ch, N_b))
    L = lib()
    wsb = L.xh_poisson_workspace_bytes(N_ch)
    ws = D.empty((wsb + 7) // 8)
    lnl = C.c_double(0.0)
    rc = L.xh_poisson_likelihood_given_background(
        float(exposure_time), N_b, D.ptr(dph), N_ch, D.ptr(dcnt), n_comp, D.ptr(dc), N_P,
        D.ptr(dcp), D.ptr(dsh), D.ptr(dbg), D.ptr(dpre), int(allow_negative), C.byref(lnl),
        D.ptr(EXPEC), D.ptr(ws), ws.numel() * 8, D.stream_ptr())
    check(rc, "xh_poisson_likelihood_given_background")
    return lnl.value, EXPEC.cpu().numpy()
