"""Drop-in for xpsi/cellmesh/integratorIQU_for_azimuthal_invariance.pyx:68-746.

`integrate(...)` keeps the reference signature (:68-97) and return value
(:737-746): ``(0, I[N_E, N_P], Q[N_E, N_P], U[N_E, N_P])`` or ``(1, None)``.
It is bound where the reference binds its Stokes integrator
(HotRegion._integratorIQU, HotRegion.py:551-572).  Supported: the blackbody
(hot_atm_ext 1, Q = U = 0) and the burst blackbody (hot_atm_ext 3, with its
Stokes Q), beaming options 0-3 (3 with the blackbody), disk blocking
(R_in < 1e6 m), the blackbody elsewhere correction (else_atm_ext 1); other
settings raise NotImplementedError.
"""
import numpy as np

from ._stokes import run


def integrate(numThreads, R, omega, r_s, inclination, cellArea,
              radialCoords_of_parallels, r_s_over_r, theta, phi, srcCellParams,
              CELL_RADIATES, correction_srcCellParams, numRays, deflection,
              cos_alpha, lag, maxDeflection, cos_gammaArray, energies, leaves,
              phases, hot_atmosphere_I, hot_atmosphere_Q, elsewhere_atmosphere,
              hot_atm_ext, else_atm_ext, beam_opt, image_order_limit=None, R_in=1e6):
    out = run(omega, inclination, cellArea, radialCoords_of_parallels, r_s_over_r, theta, phi,
              srcCellParams, CELL_RADIATES, correction_srcCellParams, deflection, cos_alpha, lag,
              maxDeflection, cos_gammaArray, energies, leaves, phases, hot_atm_ext, else_atm_ext,
              beam_opt, image_order_limit, R_in, stokes=True)
    if out is None:
        return (1, None)
    return (0,) + tuple(out)
