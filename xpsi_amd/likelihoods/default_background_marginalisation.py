"""Drop-in for xpsi/likelihoods/default_background_marginalisation.pyx.

precomputation (:38-68) and eval_marginal_likelihood (:450-734) keep the
reference signatures and return values.  The marginal over each channel's
background is integrated with QAG-GK61 to a relative tolerance of
min(epsrel, 1e-11); llzero-class outputs are drawn near llzero as the
reference does.  Components may lie on different phase grids and take
allow_negative per component, as the reference's do.
"""
import ctypes as C

import numpy as np

from .._lib import lib, check
from .. import _device as D
from ._components import allow_flags, pack


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
    flags = allow_flags(allow_negative, len(components))
    shared, n_comp, N_P, dc, dcp, dnp = pack(components, component_phases, N_ch)
    dph, dcnt = D.to_dev(phases), D.to_dev(counts)
    dsh = D.to_dev(np.asarray(phase_shifts, dtype=np.float64))
    dpre, dsup = D.to_dev(neg_sum_ln_data_factorial), D.to_dev(support)
    dbg = None if background is None else D.to_dev(background)
    STAR = D.empty((N_ch, N_b))
    bml, bmls = D.empty(N_ch), D.empty(N_ch)
    L = lib()
    wsb = L.xh_marginal_workspace_bytes(N_ch)
    ws = D.empty((wsb + 7) // 8)
    lnl = C.c_double(0.0)
    head = (float(exposure_time), N_b, D.ptr(dph), N_ch, D.ptr(dcnt), n_comp, D.ptr(dc))
    mid = (D.ptr(dsh), D.ptr(dpre), D.ptr(dsup), float(epsabs), float(epsrel), float(epsilon), float(sigmas),
           float(llzero))
    tail = (float(slim), D.ptr(dbg), C.byref(lnl), D.ptr(STAR), D.ptr(bml), D.ptr(bmls), D.ptr(ws),
            ws.numel() * 8, D.stream_ptr())
    if shared and len(set(flags)) == 1:
        rc = L.xh_eval_marginal_likelihood(*head, N_P, D.ptr(dcp), *mid, int(flags[0]), *tail)
        check(rc, "xh_eval_marginal_likelihood")
    else:
        if shared:
            # one grid, per-component flags: the same layout, the grid repeated
            dnp = D.to_dev(np.full(n_comp, N_P, dtype=np.int32), np.int32)
            dcp = D.to_dev(np.tile(dcp.cpu().numpy(), n_comp))
        af = C.cast((C.c_int * n_comp)(*[int(f) for f in flags]), C.c_void_p)
        rc = L.xh_eval_marginal_likelihood_ex(*head, D.ptr(dnp), D.ptr(dcp), *mid, af, *tail)
        check(rc, "xh_eval_marginal_likelihood_ex")
    return (lnl.value, STAR.cpu().numpy(), bml.cpu().numpy(), bmls.cpu().numpy())
