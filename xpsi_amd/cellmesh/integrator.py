"""Drop-in for xpsi/cellmesh/integrator.pyx:48-667, the integrator without
azimuthal invariance (HotRegion.symmetry False, HotRegion.py:551-572).

`integrate(...)` keeps the reference signature (:48-78) and return value
(:663-667): ``(0, flux[N_E, N_P])`` or ``(1, None)``.  Every cell takes its
own parameters from `srcCellParams` [S, S, n] (log10 T first; with beaming
log g, a, b, c, d, n_mu follow), and with `correction_srcCellParams` its own
blackbody elsewhere temperature.  Supported: the blackbody (hot_atm_ext 1)
and burst blackbody (3) with beaming options 0-3 (3 with the blackbody), with
or without the blackbody elsewhere correction (else_atm_ext 1); other
settings raise NotImplementedError.  `R_in` is accepted and ignored, as the
reference's general integrator has no disk (HotRegion.py:1166-1167 warns of
it).
"""
from ._general import run


def integrate(numThreads, R, omega, r_s, inclination, cellArea,
              radialCoords_of_parallels, r_s_over_r, theta, phi, srcCellParams,
              CELL_RADIATES, correction_srcCellParams, numRays, deflection,
              cos_alpha, lag, maxDeflection, cos_gammaArray, energies, leaves,
              phases, hot_atmosphere, elsewhere_atmosphere, hot_atm_ext,
              else_atm_ext, beam_opt, image_order_limit=None, R_in=1e6):
    out = run(omega, inclination, cellArea, radialCoords_of_parallels, r_s_over_r, theta, phi, srcCellParams,
              CELL_RADIATES, correction_srcCellParams, deflection, cos_alpha, lag, maxDeflection,
              cos_gammaArray, energies, leaves, phases, hot_atm_ext, else_atm_ext, beam_opt,
              image_order_limit, stokes=False)
    return (1, None) if out is None else (0, out[0])
