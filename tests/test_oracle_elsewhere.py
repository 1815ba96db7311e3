"""Pins the oracle's elsewhere restatement (closed mesh, time-invariant
integrator, hot-region correction) against what the mathematics fixes; the
reference holds no output for these paths.

- Closed mesh: cells of equal area summing to the star's surface, exactly
  4 pi R^2 for a sphere; ring colatitudes mirror about the equator.
- Time-invariant spectrum of a non-rotating spherical blackbody star: every
  image ray carries g^3 I(E/g) and the visible solid angle is
  pi R^2 / (1 - u) / D^2, so the integrator's sum equals
  pi R^2 g I(E/g) norm / (E keV) up to the mesh's discretisation error,
  which must fall as the mesh is refined.
- Correction: a hot region at the elsewhere temperature emits exactly
  nothing, and a corrected region equals the unclipped difference of two
  uncorrected integrations.
"""
import math

import numpy as np

from oracle import oracle as O

KB_KEV = 1.38064852e-23 / 1.60217662e-16
BB_NORM = 1.0e-7 * 5.040366110812353e22
KEV = 1.60217662e-16


def test_closed_mesh_area_and_symmetry():
    sph = O.spacetime(1.4, 12.0, 1.0, 0.5, 1.0e-9, "sphere")
    m = O.closed_mesh(sph, 16)
    assert abs(m["cellArea"] * 256 / (4.0 * math.pi * sph["R"] ** 2) - 1.0) < 1e-12
    obl = O.spacetime(1.4, 12.0, 1.0, 0.5, 600.0)
    mo = O.closed_mesh(obl, 20)
    assert np.allclose(mo["colat"][:10] + mo["colat"][::-1][:10], math.pi, rtol=0, atol=1e-14)
    assert np.all(np.diff(mo["colat"]) > 0)
    # the oblate star: equatorial radius R, polar radius smaller
    assert mo["r"][9] <= obl["R"] and mo["r"][0] < mo["r"][9]
    assert np.allclose(mo["phis"], np.linspace(-math.pi + math.pi / 20, math.pi - math.pi / 20, 20))


def test_time_invariant_sphere_closed_form():
    st = O.spacetime(1.4, 12.0, 1.0, 0.5, 1.0e-9, "sphere")
    E = np.logspace(np.log10(0.3), 1.0, 24)
    g = math.sqrt(1.0 - st["r_s"] / st["R"])
    Ep = E / g
    kT = KB_KEV * 10 ** 6.0
    ref = math.pi * st["R"] ** 2 * g * Ep ** 3 / np.expm1(Ep / kT) * BB_NORM / (E * KEV)
    devs = []
    for S in (16, 64):
        m, err = O.elsewhere_embed(st, dict(sqrt_num_cells=S, num_rays=1000))
        assert err is None
        t, spec = O.elsewhere_integrate(st, m, E, 6.0)
        assert t == 0
        devs.append(np.max(np.abs(spec / ref - 1.0)))
    assert devs[1] < 1e-4
    assert devs[1] < 0.5 * devs[0]


def _corrected_model_parts(T_hot, T_else):
    from xpsi_amd import configs
    m = configs.config1()
    sp = m["spacetime"]
    st = O.spacetime(sp["mass"], sp["radius"], sp["distance"], sp["cos_inclination"], sp["frequency"])
    hr = dict(m["hot_regions"][0], sqrt_num_cells=16, num_leaves=32, super_temperature=T_hot)
    members, err = O.hot_region_embed(st, hr, sp["frequency"])
    assert err is None
    E = np.logspace(np.log10(0.3), 1.0, 16)
    return st, hr, members, E


def test_correction_at_elsewhere_temperature_is_zero():
    st, hr, members, E = _corrected_model_parts(6.0, 6.0)
    out = O.hot_region_integrate(st, hr, members, E, else_logT=6.0)
    assert np.all(out[0] == 0.0)


def test_correction_is_unclipped_difference():
    """The corrected profile is I_hot - I_else per leaf and energy; Akima's
    slope weights are not linear in the data, so the comparison with the
    difference of two integrations holds to the interpolant's
    nonlinearity (~1e-6 here), far below what a wrong temperature or a
    missing term would leave."""
    st, hr, members, E = _corrected_model_parts(6.2, 5.9)
    corr = O.hot_region_integrate(st, hr, members, E, else_logT=5.9)[0]
    # unclipped integrations: a correction at a temperature so low that its
    # intensity underflows to zero switches the clipping off and subtracts 0
    hot = O.hot_region_integrate(st, hr, members, E, else_logT=-300.0)[0]
    m0 = dict(members[0])
    p = m0["params"].copy()
    p[..., 0] = 5.9
    m0["params"] = p
    cold = O.hot_region_integrate(st, hr, [m0], E, else_logT=-300.0)[0]
    d = hot - cold
    assert np.max(np.abs(corr - d)) < 1e-4 * np.max(np.abs(d))


def test_correction_switches_clipping_off():
    """Without correction, negative spline values are dropped, so the clipped
    integration is >= the unclipped one everywhere and differs somewhere."""
    st, hr, members, E = _corrected_model_parts(6.2, 5.9)
    unclipped = O.hot_region_integrate(st, hr, members, E, else_logT=-300.0)[0]
    clipped = O.hot_region_integrate(st, hr, members, E)[0]
    assert np.all(clipped >= unclipped)
    assert np.any(clipped > unclipped)
