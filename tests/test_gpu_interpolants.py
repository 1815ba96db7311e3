"""GPU: the phase interpolant choice (xpsi/tools/core.pyx:21-116: "Akima"
periodic by default, "Steffen", "Cubic" the periodic cubic spline) through
the drop-ins against the oracle under the same choice: the integrator with
azimuthal invariance (on the Stokes kernels), its Stokes form, the general
integrator, the energy integral, the background-marginalised and the Poisson
likelihoods (components on their own phase grids).  Signals to 1e-10 element
by element (Q, U with their 1e-5 floor), lnL to 1e-8.  The batched pipeline
keeps periodic Akima and refuses another choice."""
import math
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests.bars import LNL_ATOL, SIG_RTOL, elem_relerr
from tests.models import examples_fast_model
from tests.test_gpu_components import _case

pytestmark = pytest.mark.gpu
QU_FLOOR = 1.0e-5


def say(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(params=["Steffen", "Cubic"])
def interp(request):
    from xpsi_amd.tools.core import set_phase_interpolant
    set_phase_interpolant(request.param)
    O.set_phase_interpolant(request.param)
    yield request.param
    set_phase_interpolant("Akima")
    O.set_phase_interpolant("Akima")


@pytest.fixture(scope="module")
def ef(examples_fast):
    model = examples_fast_model(examples_fast)
    sp = model["spacetime"]
    st = O.spacetime(sp["mass"], sp["radius"], sp["distance"], sp["cos_inclination"], sp["frequency"],
                     sp.get("star_shape", "AGM_14"))
    members, err = O.hot_region_embed(st, model["hot_regions"][0], sp["frequency"])
    assert err is None
    energies = O.construct_energy_array(64, model["instrument"]["energy_edges"])
    return st, members[0], energies, model


def _geo(m):
    return (m["cellArea"], m["r"], m["r_s_over_r"], m["theta"], m["phi"])


def _rays(m):
    return (m["deflection"], m["cos_alpha"], m["lag"], m["maxDeflection"], m["cos_gamma"])


def test_invariant_integrator(ef, interp):
    from xpsi_amd.cellmesh.integrator_for_azimuthal_invariance import integrate
    st, m, energies, _ = ef
    leaves = np.linspace(0.0, 2.0 * math.pi, 64)
    phases = 2.0 * math.pi * np.linspace(0.0, 1.0, 64)
    t, ref = O.integrate(st["Omega"], st["r_s"], st["i"], *_geo(m), m["params"], m["radiates"], *_rays(m),
                         energies, leaves, phases, R=st["R"])
    status, flux = integrate(1, st["R"], st["Omega"], st["r_s"], st["i"], *_geo(m), m["params"], m["radiates"],
                             None, m["deflection"].shape[1], *_rays(m), energies, leaves, phases, (), (), 1, None, 0)
    assert t == 0 and status == 0
    err = elem_relerr(flux, ref)
    say("%s: invariant integrator rel err %.3e" % (interp, err))
    assert err < SIG_RTOL


def test_stokes_and_general(ef, interp):
    from xpsi_amd.cellmesh.integratorIQU_for_azimuthal_invariance import integrate as iqu
    from xpsi_amd.cellmesh.integratorIQU import integrate as giqu
    st, m, energies, _ = ef
    leaves = np.linspace(0.0, 2.0 * math.pi, 64)
    phases = 2.0 * math.pi * np.linspace(0.0, 1.0, 64)
    P = m["params"].copy()
    P[..., 0] += 0.1 * np.sin(m["phi"])
    ref = O.integrate_stokes(st["Omega"], st["i"], *_geo(m), m["params"], m["radiates"], *_rays(m), energies,
                             leaves, phases, atm_ext=3)
    got = iqu(1, st["R"], st["Omega"], st["r_s"], st["i"], *_geo(m), m["params"], m["radiates"], None,
              m["deflection"].shape[1], *_rays(m), energies, leaves, phases, (), (), (), 3, None, 0)
    gref = O.integrate_general_stokes(st["Omega"], st["i"], *_geo(m), P, m["radiates"], *_rays(m), energies,
                                      leaves, phases, atm_ext=3)
    ggot = giqu(1, st["R"], st["Omega"], st["r_s"], st["i"], *_geo(m), P, m["radiates"], None,
                m["deflection"].shape[1], *_rays(m), energies, leaves, phases, (), (), (), 3, None, 0)
    for tag, a, b in (("Stokes", got, ref), ("general Stokes", ggot, gref)):
        assert a[0] == 0 and b[0] == 0
        for nm, x, y in zip("IQU", a[1:], b[1:]):
            err = elem_relerr(x, y) if nm == "I" else elem_relerr(x, y, floor=QU_FLOOR)
            say("%s: %s %s rel err %.3e" % (interp, tag, nm, err))
            assert err < SIG_RTOL


def test_energy_integrator(ef, interp):
    from xpsi_amd.tools import energy_integrator
    st, m, energies, model = ef
    rng = np.random.default_rng(3)
    sig = np.abs(rng.normal(1.0, 0.3, (len(energies), 24))) * energies[:, None] ** -2
    edges = np.log10(model["instrument"]["energy_edges"])
    ref = O.energy_integrator(sig, np.log10(energies), edges)
    got = energy_integrator(1, sig, np.log10(energies), edges)
    err = elem_relerr(got, ref)
    say("%s: energy integral rel err %.3e" % (interp, err))
    assert err < SIG_RTOL


def test_likelihoods(interp):
    from xpsi_amd.likelihoods import default_background_marginalisation as G
    from xpsi_amd.likelihoods._poisson_likelihood_given_background import poisson_likelihood_given_background as P
    phases, counts, comps, grids, shifts, bg = _case(seed=11)
    pre = O.precomputation(counts.astype(np.int32))
    sup = np.zeros((counts.shape[0], 2))
    sup[:, 1] = -1.0
    out = G.eval_marginal_likelihood(100.0, phases, counts, comps, grids, shifts, pre, sup, 1000, 0.0, 1.0e-8,
                                     1.0e-3, 10.0, -1.0e90, allow_negative=[True, False])
    ref = O.eval_marginal_likelihood(100.0, phases, counts, comps, grids, shifts, pre, sup,
                                     allow_negative=[True, False])
    say("%s: marginal lnL gpu %.10f oracle %.10f" % (interp, out[0], ref[0]))
    assert abs(out[0] - ref[0]) < LNL_ATOL
    assert elem_relerr(out[1], ref[1]) < SIG_RTOL
    got = P(100.0, phases, counts, comps, grids, shifts, bg)
    pref = O.poisson_likelihood_given_background(100.0, phases, counts, comps, grids, shifts, bg)
    say("%s: Poisson lnL gpu %.10f oracle %.10f" % (interp, got[0], pref[0]))
    assert abs(got[0] - pref[0]) < LNL_ATOL
    assert elem_relerr(got[1], pref[1]) < SIG_RTOL


def test_pipeline_refuses_other_interpolants(interp):
    from xpsi_amd import configs
    from xpsi_amd._lib import XHError
    from xpsi_amd.pipeline import LikelihoodPipeline
    with pytest.raises(XHError):
        LikelihoodPipeline(configs.config1(), max_batch=1)
