"""The phase interpolant choice of xpsi/tools/core.pyx:21-116.

`set_phase_interpolant("Akima" | "Steffen" | "Cubic")` selects, for every
later drop-in call in the process, the interpolant of the integrators' phase
profiles, of the energy integral and of the likelihoods' phase-bin integrals:
periodic Akima (the default), Steffen, or the periodic cubic spline.  The
integrator with azimuthal invariance then runs on the Stokes integrator's
kernels (blackbody atmospheres); the batched pipeline keeps periodic Akima and
refuses another choice at creation.
"""
from .._lib import lib, check

__interpolants__ = {"Akima": 0, "Steffen": 1, "Cubic": 2}


def set_phase_interpolant(interpolant):
    if not isinstance(interpolant, str):
        raise TypeError("Interpolant declaration must be a string.")
    if interpolant not in __interpolants__:
        raise ValueError("Invalid interpolant name: one of %s." % ", ".join(__interpolants__))
    check(lib().xh_set_phase_interpolant(__interpolants__[interpolant]), "xh_set_phase_interpolant")


def get_phase_interpolant():
    k = lib().xh_get_phase_interpolant()
    for name, v in __interpolants__.items():
        if v == k:
            return name
    raise RuntimeError("unknown phase interpolant %d" % k)
