"""Drop-in for xpsi/cellmesh/integratorIQU.pyx:51-717, the Stokes form of the
integrator without azimuthal invariance (HotRegion._integratorIQU with
symmetry False, HotRegion.py:569-570).

`integrate(...)` keeps the reference signature (:51-80) and return value (:715-717):
``(0, I[N_E, N_P], Q[N_E, N_P], U[N_E, N_P])`` or ``(1, None)``.  Q and U
are the atmosphere's Q at the interpolated mu eta rotated by twice the
polarisation angle, which the reference forms from its leaf loop's last values
(:593-596), so one angle per parallel and image.  Supported: as the plain
general integrator (blackbody: Q = U = 0; burst blackbody; beaming; the
blackbody elsewhere correction per cell).
"""
from ._general import run


def integrate(numThreads, R, omega, r_s, inclination, cellArea,
              radialCoords_of_parallels, r_s_over_r, theta, phi, srcCellParams,
              CELL_RADIATES, correction_srcCellParams, numRays, deflection,
              cos_alpha, lag, maxDeflection, cos_gammaArray, energies, leaves,
              phases, hot_atmosphere_I, hot_atmosphere_Q, elsewhere_atmosphere,
              hot_atm_ext, else_atm_ext, beam_opt, image_order_limit=None, R_in=1e6):
    out = run(omega, inclination, cellArea, radialCoords_of_parallels, r_s_over_r, theta, phi, srcCellParams,
              CELL_RADIATES, correction_srcCellParams, deflection, cos_alpha, lag, maxDeflection,
              cos_gammaArray, energies, leaves, phases, hot_atm_ext, else_atm_ext, beam_opt,
              image_order_limit, stokes=True)
    if out is None:
        return (1, None)
    return (0,) + tuple(out)
