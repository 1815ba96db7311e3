"""GPU: the integrator without azimuthal invariance (xpsi/cellmesh/
integrator.pyx:48-667) through its drop-in against the oracle, signals to
1e-10 element by element.  Cases: the examples_fast spot with a per-cell
temperature map (up to three image orders); the same with a per-cell
elsewhere correction; a spot low on the far hemisphere, hidden for part of
the rotation (the padded stretches of phase, redshift and GEOM); and a ragged
pole-centred cap with a hot sector (41 leaves, 100 energies, 100 phases:
part-filled energy and phase tiles).  The Stokes form
(xpsi/cellmesh/integratorIQU.pyx:51-717) with the burst atmosphere, I, Q and
U, and the beaming options."""
import math
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests.bars import SIG_RTOL, elem_relerr
from tests.models import examples_fast_model

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


def _tmap(m, seed, amp=0.15):
    rng = np.random.default_rng(seed)
    P = m["params"].copy()
    a, b = rng.uniform(-1.0, 1.0, 2)
    P[..., 0] = P[..., 0] + amp * np.sin(m["phi"] + a) * np.cos(m["theta"] * 3.0 + b)
    return P


def _compare(st, m, energies, N_L, N_P, params, corr=None, image_order_limit=None, name=""):
    from xpsi_amd.cellmesh.integrator import integrate
    leaves = np.linspace(0.0, 2.0 * math.pi, N_L)
    phases = 2.0 * math.pi * np.linspace(0.0, 1.0, N_P)
    t, ref = O.integrate_general(st["Omega"], st["i"], m["cellArea"], m["r"], m["r_s_over_r"], m["theta"],
                                 m["phi"], params, m["radiates"], m["deflection"], m["cos_alpha"], m["lag"],
                                 m["maxDeflection"], m["cos_gamma"], energies, leaves, phases,
                                 image_order_limit=image_order_limit, correction=corr)
    assert t == 0
    say("%s: S=%d N_R=%d N_E=%d N_L=%d N_P=%d" % (name, m["S"], m["deflection"].shape[1], len(energies),
                                                  N_L, N_P))
    status, flux = integrate(1, st["R"], st["Omega"], st["r_s"], st["i"], m["cellArea"], m["r"],
                             m["r_s_over_r"], m["theta"], m["phi"], params, m["radiates"], corr,
                             m["deflection"].shape[1], m["deflection"], m["cos_alpha"], m["lag"],
                             m["maxDeflection"], m["cos_gamma"], energies, leaves, phases, (), (), 1,
                             None if corr is None else 1, 0, image_order_limit=image_order_limit)
    assert status == 0
    err = elem_relerr(flux, ref)
    say("%s: rel err %.3e" % (name, err))
    assert err < SIG_RTOL
    return flux, ref


@pytest.fixture(scope="module")
def ef(examples_fast):
    model = examples_fast_model(examples_fast)
    sp = model["spacetime"]
    st = O.spacetime(sp["mass"], sp["radius"], sp["distance"], sp["cos_inclination"], sp["frequency"],
                     sp.get("star_shape", "AGM_14"))
    members, err = O.hot_region_embed(st, model["hot_regions"][0], sp["frequency"])
    assert err is None
    energies = O.construct_energy_array(128, model["instrument"]["energy_edges"])
    return st, members[0], energies


def test_general_temperature_map(ef):
    st, m, energies = ef
    _compare(st, m, energies, 64, 64, _tmap(m, 1), image_order_limit=3, name="examples_fast, T map")


def test_general_temperature_map_correction(ef):
    st, m, energies = ef
    corr = np.ones_like(m["params"])
    corr[..., 0] = 5.9 + 0.05 * np.cos(m["phi"])
    _compare(st, m, energies, 64, 64, _tmap(m, 2), corr=corr, name="examples_fast, T map, correction")


def test_general_hidden_stretches():
    st = O.spacetime(1.6, 13.0, 1.0, math.cos(1.1), 400.0, "AGM_14")
    hr = dict(super_colatitude=2.3, super_radius=0.5, super_temperature=6.25, phase_shift=0.0,
              sqrt_num_cells=32, min_sqrt_num_cells=10, max_sqrt_num_cells=64, num_rays=256,
              num_leaves=64)
    members, err = O.hot_region_embed(st, hr, st["f"])
    assert err is None
    m = members[0]
    energies = np.logspace(-0.7, 1.0, 64)
    flux, ref = _compare(st, m, energies, 64, 48, _tmap(m, 3), name="far-hemisphere spot")
    # the spot is hidden for part of the rotation
    assert np.min(ref[20]) < 1e-3 * np.max(ref[20])


def test_general_ragged_sector():
    st = O.spacetime(1.4, 12.0, 1.0, math.cos(1.0), 300.0, "AGM_14")
    hr = dict(super_colatitude=0.0, super_radius=0.6, super_temperature=6.2, phase_shift=0.0,
              sqrt_num_cells=24, min_sqrt_num_cells=10, max_sqrt_num_cells=100, num_rays=300,
              num_leaves=41)
    members, err = O.hot_region_embed(st, hr, st["f"])
    assert err is None
    m = members[0]
    S = m["phi"].shape[1]
    P = m["params"].copy()
    P[..., 0] = 6.0
    P[:, S // 3: S // 3 + 4, 0] = 6.35
    energies = np.logspace(-0.5, 1.0, 100)
    _compare(st, m, energies, 41, 100, P, name="ragged sector")


QU_FLOOR = 1.0e-5   # Q and U change sign across the spot: see test_gpu_stokes.py


def _compare_stokes(st, m, energies, N_L, N_P, params, atm, beam=0, corr=None, name=""):
    from xpsi_amd.cellmesh.integratorIQU import integrate
    leaves = np.linspace(0.0, 2.0 * math.pi, N_L)
    phases = 2.0 * math.pi * np.linspace(0.0, 1.0, N_P)
    ref = O.integrate_general_stokes(st["Omega"], st["i"], m["cellArea"], m["r"], m["r_s_over_r"], m["theta"],
                                     m["phi"], params, m["radiates"], m["deflection"], m["cos_alpha"], m["lag"],
                                     m["maxDeflection"], m["cos_gamma"], energies, leaves, phases, atm_ext=atm,
                                     beam_opt=beam, correction=corr)
    assert ref[0] == 0
    say("%s: S=%d N_E=%d N_L=%d N_P=%d" % (name, m["S"], len(energies), N_L, N_P))
    got = integrate(1, st["R"], st["Omega"], st["r_s"], st["i"], m["cellArea"], m["r"], m["r_s_over_r"],
                    m["theta"], m["phi"], params, m["radiates"], corr, m["deflection"].shape[1],
                    m["deflection"], m["cos_alpha"], m["lag"], m["maxDeflection"], m["cos_gamma"], energies,
                    leaves, phases, (), (), (), atm, None if corr is None else 1, beam)
    assert got[0] == 0
    for nm, a, b in zip("IQU", got[1:], ref[1:]):
        err = elem_relerr(a, b) if nm == "I" else elem_relerr(a, b, floor=QU_FLOOR)
        say("  %s rel err %.3e (max |%s| %.3e)" % (nm, err, nm, np.max(np.abs(b))))
        assert err < SIG_RTOL


def test_general_stokes_burst_temperature_map(ef):
    st, m, energies = ef
    _compare_stokes(st, m, energies, 64, 64, _tmap(m, 4), 3, name="general Stokes, burst, T map")


def test_general_stokes_burst_correction_hidden():
    st = O.spacetime(1.6, 13.0, 1.0, math.cos(1.1), 400.0, "AGM_14")
    hr = dict(super_colatitude=2.3, super_radius=0.5, super_temperature=6.25, phase_shift=0.0,
              sqrt_num_cells=32, min_sqrt_num_cells=10, max_sqrt_num_cells=64, num_rays=256,
              num_leaves=64)
    members, err = O.hot_region_embed(st, hr, st["f"])
    assert err is None
    m = members[0]
    corr = np.ones_like(m["params"])
    corr[..., 0] = 5.9
    _compare_stokes(st, m, np.logspace(-0.7, 1.0, 40), 41, 70, _tmap(m, 5), 3, corr=corr,
                    name="general Stokes, hidden stretches, correction, ragged")


@pytest.mark.parametrize("beam", [1, 2, 3])
def test_general_beaming(ef, beam):
    from xpsi_amd.cellmesh.integrator import integrate
    st, m, energies = ef
    P = np.zeros(m["params"].shape[:2] + (7,))
    P[..., :2] = _tmap(m, 6)
    P[..., 2], P[..., 3], P[..., 4], P[..., 5], P[..., 6] = 0.4, -0.3, 0.2, 0.1, 50
    leaves = np.linspace(0.0, 2.0 * math.pi, 64)
    phases = 2.0 * math.pi * np.linspace(0.0, 1.0, 64)
    t, ref = O.integrate_general_stokes(st["Omega"], st["i"], m["cellArea"], m["r"], m["r_s_over_r"], m["theta"],
                                        m["phi"], P, m["radiates"], m["deflection"], m["cos_alpha"], m["lag"],
                                        m["maxDeflection"], m["cos_gamma"], energies, leaves, phases,
                                        beam_opt=beam, stokes=False)
    assert t == 0
    status, flux = integrate(1, st["R"], st["Omega"], st["r_s"], st["i"], m["cellArea"], m["r"],
                             m["r_s_over_r"], m["theta"], m["phi"], P, m["radiates"], None,
                             m["deflection"].shape[1], m["deflection"], m["cos_alpha"], m["lag"],
                             m["maxDeflection"], m["cos_gamma"], energies, leaves, phases, (), (), 1, None, beam)
    assert status == 0
    err = elem_relerr(flux, ref)
    say("general, beaming option %d: rel err %.3e" % (beam, err))
    assert err < SIG_RTOL
