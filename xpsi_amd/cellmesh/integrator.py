"""Drop-in for xpsi/cellmesh/integrator.pyx:48-667, the integrator without
azimuthal invariance (HotRegion.symmetry False, HotRegion.py:551-572).

`integrate(...)` keeps the reference signature (:48-78) and return value
(:663-667): ``(0, flux[N_E, N_P])`` or ``(1, None)``.  Every cell takes its
own parameters from `srcCellParams` [S, S, 2] (log10 T first), and with
`correction_srcCellParams` its own blackbody elsewhere temperature.
Supported: the blackbody (hot_atm_ext 1) without beaming, with or without the
blackbody elsewhere correction (else_atm_ext 1); other settings raise
NotImplementedError.  `R_in` is accepted and ignored, as the reference's
general integrator has no disk (HotRegion.py:1166-1167 warns of it).
"""
import ctypes as C

import numpy as np

from .._lib import lib, check
from .. import _device as D


def integrate(numThreads, R, omega, r_s, inclination, cellArea,
              radialCoords_of_parallels, r_s_over_r, theta, phi, srcCellParams,
              CELL_RADIATES, correction_srcCellParams, numRays, deflection,
              cos_alpha, lag, maxDeflection, cos_gammaArray, energies, leaves,
              phases, hot_atmosphere, elsewhere_atmosphere, hot_atm_ext,
              else_atm_ext, beam_opt, image_order_limit=None, R_in=1e6):
    if int(hot_atm_ext) != 1 or int(beam_opt) != 0:
        raise NotImplementedError("the general integrator supports the blackbody without beaming")
    corr = None
    if correction_srcCellParams is not None:
        if int(else_atm_ext) != 1:
            raise NotImplementedError("elsewhere correction: blackbody only")
        corr = D.to_dev(np.ascontiguousarray(correction_srcCellParams, dtype=np.float64))
    S = int(np.shape(cellArea)[0])
    if np.shape(cellArea) != (S, S) or np.shape(srcCellParams)[:2] != (S, S):
        raise ValueError("cellArea and srcCellParams must be [S, S] and [S, S, n]")
    params = np.ascontiguousarray(np.asarray(srcCellParams, dtype=np.float64)[..., :2])
    if params.shape[2] < 2:
        params = np.concatenate([params, np.zeros((S, S, 1))], axis=2)
    N_R = int(np.shape(deflection)[1])
    N_E, N_L, N_P = len(energies), len(leaves), len(phases)
    dev = [D.to_dev(a) for a in (cellArea, radialCoords_of_parallels, r_s_over_r, theta, phi, params)]
    rad = D.to_dev(CELL_RADIATES, np.int32)
    rays = [D.to_dev(a) for a in (deflection, cos_alpha, lag, maxDeflection,
                                   cos_gammaArray, energies, leaves, phases)]
    flux = D.empty((N_E, N_P))
    L = lib()
    wsb = L.xh_integrate_general_workspace_bytes(S, N_L)
    ws = D.empty((wsb + 7) // 8)
    status = C.c_int(0)
    rc = L.xh_integrate_general(float(omega), float(inclination), S, D.ptr(dev[0]), D.ptr(dev[1]),
                                D.ptr(dev[2]), D.ptr(dev[3]), D.ptr(dev[4]), D.ptr(dev[5]), D.ptr(rad), N_R,
                                D.ptr(rays[0]), D.ptr(rays[1]), D.ptr(rays[2]), D.ptr(rays[3]),
                                D.ptr(rays[4]), N_E, D.ptr(rays[5]), N_L, D.ptr(rays[6]), N_P,
                                D.ptr(rays[7]), int(image_order_limit or 0), D.ptr(corr), D.ptr(flux),
                                D.ptr(ws), ws.numel() * 8, C.byref(status), D.stream_ptr())
    check(rc, "xh_integrate_general")
    if status.value != 0:
        return (1, None)
    return (0, flux.cpu().numpy())
