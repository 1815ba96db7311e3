"""Shared host side of the integrator without azimuthal invariance
(xh_integrate_general): the plain drop-in (xpsi/cellmesh/integrator.pyx) and
its Stokes form (xpsi/cellmesh/integratorIQU.pyx) both call it."""
import ctypes as C

import numpy as np

from .._lib import lib, check
from .. import _device as D


def run(omega, inclination, cellArea, radial, r_s_over_r, theta, phi, srcCellParams, CELL_RADIATES,
        correction_srcCellParams, deflection, cos_alpha, lag, maxDeflection, cos_gammaArray, energies,
        leaves, phases, hot_atm_ext, else_atm_ext, beam_opt, image_order_limit, stokes):
    """Returns [I] (stokes False) or [I, Q, U] as [N_E, N_P] arrays, or None
    for the reference's ERROR status."""
    atm, beam = int(hot_atm_ext), int(beam_opt)
    if atm not in (1, 3):
        raise NotImplementedError("the general integrator supports the blackbody (1) and burst blackbody (3)")
    if beam not in (0, 1, 2, 3) or (beam == 3 and atm != 1):
        raise NotImplementedError("beaming options 0-3 (3 with the blackbody)")
    S = int(np.shape(cellArea)[0])
    params = np.ascontiguousarray(srcCellParams, dtype=np.float64)
    if np.shape(cellArea) != (S, S) or params.ndim != 3 or params.shape[:2] != (S, S):
        raise ValueError("cellArea and srcCellParams must be [S, S] and [S, S, n]")
    if params.shape[2] < 2:
        params = np.concatenate([params, np.zeros((S, S, 2 - params.shape[2]))], axis=2)
    if beam != 0 and params.shape[2] < 7:
        raise ValueError("beaming needs 7 parameters per cell (log T, log g, a, b, c, d, n_mu)")
    corr = None
    if correction_srcCellParams is not None:
        if int(else_atm_ext) != 1:
            raise NotImplementedError("elsewhere correction: blackbody only")
        corr = D.to_dev(np.ascontiguousarray(correction_srcCellParams, dtype=np.float64)[..., :2])
    N_R = int(np.shape(deflection)[1])
    N_E, N_L, N_P = len(energies), len(leaves), len(phases)
    dev = [D.to_dev(a) for a in (cellArea, radial, r_s_over_r, theta, phi, params)]
    rad = D.to_dev(CELL_RADIATES, np.int32)
    rays = [D.to_dev(a) for a in (deflection, cos_alpha, lag, maxDeflection,
                                   cos_gammaArray, energies, leaves, phases)]
    fI = D.empty((N_E, N_P))
    fQ = D.empty((N_E, N_P)) if stokes else None
    fU = D.empty((N_E, N_P)) if stokes else None
    L = lib()
    wsb = L.xh_integrate_general_workspace_bytes(S, N_L)
    ws = D.empty((wsb + 7) // 8)
    status = C.c_int(0)
    rc = L.xh_integrate_general(float(omega), float(inclination), S, D.ptr(dev[0]), D.ptr(dev[1]),
                                D.ptr(dev[2]), D.ptr(dev[3]), D.ptr(dev[4]), D.ptr(dev[5]), params.shape[2],
                                D.ptr(rad), N_R, D.ptr(rays[0]), D.ptr(rays[1]), D.ptr(rays[2]),
                                D.ptr(rays[3]), D.ptr(rays[4]), N_E, D.ptr(rays[5]), N_L, D.ptr(rays[6]), N_P,
                                D.ptr(rays[7]), int(image_order_limit or 0), atm, beam, D.ptr(corr), D.ptr(fI),
                                D.ptr(fQ), D.ptr(fU), D.ptr(ws), ws.numel() * 8, C.byref(status),
                                D.stream_ptr())
    check(rc, "xh_integrate_general")
    if status.value != 0:
        return None
    return [f.cpu().numpy() for f in ((fI, fQ, fU) if stokes else (fI,))]
