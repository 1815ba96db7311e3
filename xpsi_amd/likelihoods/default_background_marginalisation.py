"""Drop-in for xpsi/likelihoods/default_background_marginalisation.pyx.

precomputation (:38-68) and eval_marginal_likelihood (:450-734) keep the
reference signatures and return values.  The marginal over each channel's
background is integrated with QAG-GK61 to a relative tolerance of
min(epsrel, 1e-11); llzero-class outputs are drawn near llzero as the
reference does.  Components must share one phase grid.
"""
import ctypes as C

import numpy as np

from .._lib import lib, check
from .. import _device as D


def precomputation(data):
    data = np.ascontiguousarray(data, dtype=np.int32)
    out = np.zeros(data.shape[0])
    check(lib().xh_precomputation(data.shape[0], data.shape[1],
                                  data.ctypes.data, out.ctypes.data), "xh_precomputation")
    return out


def eval_marginal_likelihood(exposure_time, phases, counts, components,
                             component_phases, phase_shifts,
                             neg_sum_ln_data_factorial, support,
                             workspace_intervals, epsabs, epsrel, epsilon,
                             sigmas, llzero, allow_negative=False, slim=20.0,
                             background=None):
    counts = np.ascontiguousarray(counts, dtype=np.float64)
    N_ch, N_b = counts.shape
    support = np.ascontiguousarray(support, dtype=np.float64)
    if support.shape[0] != N_ch:
        raise TypeError('The number of energy channels in the background support '
                        'does not match to that of the data.')
    for i in range(N_ch):
        if support[i, 1] == 0:
            raise TypeError('Background upper limit cannot be set to 0.')
        if support[i, 1] > 0 and support[i, 1] - support[i, 0] < 0:
            raise TypeError('Background upper limit must be higher than the lower limit.')
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
    dc = D.to_dev(comps)
    dph, dcnt = D.to_dev(phases), D.to_dev(counts)
    dcp, dsh = D.to_dev(ph0), D.to_dev(np.asarray(phase_shifts, dtype=np.float64))
    dpre, dsup = D.to_dev(neg_sum_ln_data_factorial), D.to_dev(support)
    dbg = None if background is None else D.to_dev(background)
    STAR = D.empty((N_ch, N_b))
    bml, bmls = D.empty(N_ch), D.empty(N_ch)
    L = lib()
    wsb = L.xh_marginal_workspace_bytes(N_ch)
    ws = D.empty((wsb + 7) // 8)
    lnl = C.c_double(0.0)
    rc = L.xh_eval_marginal_likelihood(
        float(exposure_time), N_b, D.ptr(dph), N_ch, D.ptr(dcnt), n_comp, D.ptr(dc),
        N_P, D.ptr(dcp), D.ptr(dsh), D.ptr(dpre), D.ptr(dsup), float(epsabs),
        float(epsrel), float(epsilon), float(sigmas), float(llzero),
        int(allow_negative), float(slim), D.ptr(dbg), C.byref(lnl), D.ptr(STAR),
        D.ptr(bml), D.ptr(bmls), D.ptr(ws), ws.numel() * 8, D.stream_ptr())
    check(rc, "xh_eval_marginal_likelihood")
    return (lnl.value, STAR.cpu().numpy(), bml.cpu().numpy(), bmls.cpu().numpy())
