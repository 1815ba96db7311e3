"""Drop-in for xpsi/cellmesh/integrator_for_time_invariance.pyx:59-338.

`integrate(...)` keeps the reference signature (:59-80) and return value
(:325-338): ``(0, flux[N_E])`` or ``(1, None)``.  It is bound where the
reference binds it (Elsewhere.integrate, xpsi/Elsewhere.py:386-441).
Supported: the blackbody elsewhere atmosphere (atm_ext 1); other settings
raise NotImplementedError.
"""
import ctypes as C

import numpy as np

from .._lib import lib, check
from .. import _device as D


def integrate(numThreads, R, omega, r_s, inclination, sqrt_numPix, cellArea,
              radialCoords_of_parallels, r_s_over_r, theta, phi, srcCellParams,
              numRays, deflection, cos_alpha, maxDeflection, cos_gammaArray,
              energies, atmosphere, atm_ext, image_order_limit=None, *args):
    if int(atm_ext) != 1:
        raise NotImplementedError("the blackbody elsewhere atmosphere only")
    S = int(sqrt_numPix)
    N_R = int(np.shape(deflection)[1])
    N_E = len(energies)
    params = np.ascontiguousarray(srcCellParams, dtype=np.float64)
    colat = np.ascontiguousarray(np.asarray(theta, dtype=np.float64)[:, 0])
    dev = [D.to_dev(a) for a in (radialCoords_of_parallels, colat, phi, params, deflection,
                                  cos_alpha, maxDeflection, cos_gammaArray, energies)]
    flux = D.empty((N_E,))
    L = lib()
    ws = D.empty((L.xh_integrate_time_invariant_workspace_bytes(S, N_E) + 7) // 8)
    status = C.c_int(0)
    rc = L.xh_integrate_time_invariant(float(omega), float(r_s), float(inclination), S,
                                       float(cellArea), D.ptr(dev[0]), D.ptr(dev[1]),
                                       D.ptr(dev[2]), D.ptr(dev[3]), int(params.shape[-1]), N_R,
                                       D.ptr(dev[4]), D.ptr(dev[5]), D.ptr(dev[6]),
                                       D.ptr(dev[7]), N_E, D.ptr(dev[8]),
                                       int(image_order_limit or 0), D.ptr(flux), D.ptr(ws),
                                       ws.numel() * 8, C.byref(status), D.stream_ptr())
    check(rc, "xh_integrate_time_invariant")
    if status.value != 0:
        return (1, None)
    return (0, flux.cpu().numpy())
