"""Drop-in for xpsi/cellmesh/integrator_for_azimuthal_invariance.pyx:70-665.

`integrate(...)` keeps the reference signature (:70-98) and return value
(:657-665): ``(0, flux[N_E, N_P])`` or ``(1, None)``.  It is bound where the
reference binds its integrator (HotRegion._integrator, HotRegion.py:551-572).
Supported: the blackbody (hot_atm_ext 1) and four-dimensional table
(hot_atm_ext 2, `hot_atmosphere` = (logT, logg, mu, logE, buf)) atmospheres,
with or without the blackbody elsewhere correction (else_atm_ext 1) at a
temperature uniform over the cells (Elsewhere._compute_cellParamVecs).  The
burst blackbody (hot_atm_ext 3), beaming (beam_opt 1-3), disk blocking
(R_in < 1e6 m) and a phase interpolant other than periodic Akima
(xpsi_amd.tools.core.set_phase_interpolant) go through the Stokes
integrator's kernels (intensity only), for the blackbody atmospheres; other
settings raise NotImplementedError.
"""
import ctypes as C

import numpy as np

from .._lib import lib, check
from .. import _device as D
from ._stokes import run


def integrate(numThreads, R, omega, r_s, inclination, cellArea,
              radialCoords_of_parallels, r_s_over_r, theta, phi, srcCellParams,
              CELL_RADIATES, correction_srcCellParams, numRays, deflection,
              cos_alpha, lag, maxDeflection, cos_gammaArray, energies, leaves,
              phases, hot_atmosphere, elsewhere_atmosphere, hot_atm_ext,
              else_atm_ext, beam_opt, image_order_limit=None, R_in=1e6):
    other_interp = lib().xh_get_phase_interpolant() != 0
    if int(hot_atm_ext) == 3 or int(beam_opt) != 0 or R_in < 1e6 or other_interp:
        if int(hot_atm_ext) not in (1, 3):
            raise NotImplementedError("beaming, disk blocking and phase interpolants other than Akima: "
                                      "blackbody atmospheres only")
        out = run(omega, inclination, cellArea, radialCoords_of_parallels, r_s_over_r, theta, phi,
                  srcCellParams, CELL_RADIATES, correction_srcCellParams, deflection, cos_alpha, lag,
                  maxDeflection, cos_gammaArray, energies, leaves, phases, hot_atm_ext, else_atm_ext,
                  beam_opt, image_order_limit, R_in, stokes=False)
        return (1, None) if out is None else (0, out[0])
    correct, else_logT = 0, 0.0
    if correction_srcCellParams is not None:
        corr = np.asarray(correction_srcCellParams, dtype=np.float64)
        if int(else_atm_ext) != 1 or not np.all(corr[..., 0] == corr.flat[0]):
            raise NotImplementedError("elsewhere correction: blackbody at one temperature")
        correct, else_logT = 1, float(corr.flat[0])
    if int(hot_atm_ext) not in (1, 2):
        raise NotImplementedError("blackbody, burst blackbody or 4-D table atmospheres")
    S = int(np.shape(cellArea)[0])
    N_R = int(np.shape(deflection)[1])
    N_E, N_L, N_P = len(energies), len(leaves), len(phases)
    dev = [D.to_dev(a) for a in (cellArea, radialCoords_of_parallels, r_s_over_r,
                                  theta, phi, srcCellParams)]
    rad = D.to_dev(CELL_RADIATES, np.int32)
    rays = [D.to_dev(a) for a in (deflection, cos_alpha, lag, maxDeflection,
                                   cos_gammaArray, energies, leaves, phases)]
    flux = D.empty((N_E, N_P))
    axes = nax = tab = None
    if int(hot_atm_ext) == 2:
        axes = D.to_dev(np.concatenate([np.asarray(a, dtype=np.float64) for a in hot_atmosphere[:4]]))
        nax = D.to_dev(np.array([len(a) for a in hot_atmosphere[:4]], dtype=np.int32), np.int32)
        tab = D.to_dev(hot_atmosphere[4])
    L = lib()
    wsb = L.xh_integrate_workspace_bytes(S, N_R, N_L, N_E, N_P)
    ws = D.empty((wsb + 7) // 8)
    status = C.c_int(0)
    rc = L.xh_integrate(float(omega), float(r_s), float(inclination), S,
                        D.ptr(dev[0]), D.ptr(dev[1]), D.ptr(dev[2]), D.ptr(dev[3]),
                        D.ptr(dev[4]), D.ptr(dev[5]), D.ptr(rad), N_R,
                        D.ptr(rays[0]), D.ptr(rays[1]), D.ptr(rays[2]),
                        D.ptr(rays[3]), D.ptr(rays[4]), N_E, D.ptr(rays[5]),
                        N_L, D.ptr(rays[6]), N_P, D.ptr(rays[7]), int(hot_atm_ext),
                        D.ptr(axes), D.ptr(nax), D.ptr(tab),
                        int(image_order_limit or 0), correct, else_logT, D.ptr(flux), D.ptr(ws),
                        ws.numel() * 8, C.byref(status), D.stream_ptr())
    check(rc, "xh_integrate")
    if status.value != 0:
        return (1, None)
    return (0, flux.cpu().numpy())
