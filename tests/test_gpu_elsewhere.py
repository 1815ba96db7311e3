"""GPU parity of the elsewhere component (closed mesh, time-invariant
spectrum, hot-region correction) and interstellar absorption: the large
workload's model at parity-test size against the oracle, FP64 bars (signals
1e-10 relative, lnL 1e-8 absolute), a batch of its prior draws, and the
precision form against FP64 (|dlnL| < 1e-3)."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests.bars import elem_relerr, norm_relerr

pytestmark = pytest.mark.gpu

SIG_RTOL = 1.0e-10
LNL_ATOL = 1.0e-8
LLZERO = -1.0e90


def say(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


relerr = elem_relerr


def test_elsewhere_spectrum_signals_and_lnl():
    from xpsi_amd import configs
    from xpsi_amd.pipeline import LikelihoodPipeline, params_from_model
    model = configs.config5("test")
    lnl_ref, parts = O.loglikelihood(model, return_parts=True)
    sp = model["spacetime"]
    st = O.spacetime(sp["mass"], sp["radius"], sp["distance"], sp["cos_inclination"], sp["frequency"])
    em, err = O.elsewhere_embed(st, model["elsewhere"])
    assert err is None
    t, spec_ref = O.elsewhere_integrate(st, em, parts["energies"], 5.8)
    assert t == 0
    say("config5-test: creating pipeline (4 hot members + elsewhere)")
    pl = LikelihoodPipeline(model, max_batch=1)
    lnl = pl.evaluate([params_from_model(model)])[0]
    spec = pl.fetch("else_spectrum", 0, 0, len(spec_ref))[:len(spec_ref)]
    e_spec = relerr(spec, spec_ref)
    say("config5-test: elsewhere spectrum rel err %.3e" % e_spec)
    assert e_spec < SIG_RTOL
    for r in range(2):
        fo = pl.fetch("folded", 0, r, pl.N_ch * pl.N_P)[:pl.N_ch * pl.N_P].reshape(pl.N_ch, pl.N_P)
        e = relerr(fo, parts["signals"][r])
        say("config5-test: region %d folded rel err %.3e" % (r, e))
        assert e < SIG_RTOL
    say("config5-test: lnL gpu %.10f oracle %.10f" % (lnl, lnl_ref))
    assert abs(lnl - lnl_ref) < LNL_ATOL
    pl.close()


def _draw_model(model, d):
    m = dict(model)
    m["spacetime"] = dict(model["spacetime"], mass=d["mass"], radius=d["radius"],
                          distance=d["distance"], cos_inclination=d["cos_inclination"])
    m["hot_regions"] = [dict(h, **hp) for h, hp in zip(model["hot_regions"], d["hot_regions"])]
    m["elsewhere"] = dict(model["elsewhere"], elsewhere_temperature=d["elsewhere_temperature"])
    return m


def test_elsewhere_batch_of_draws():
    from xpsi_amd import configs
    from xpsi_amd.pipeline import LikelihoodPipeline
    model = configs.config5("test")
    draws = [d for d in configs.config5_draws(64)
             if configs.prior_accepts(d["mass"], d["radius"], 300.0)][:8]
    ref = [O.loglikelihood(_draw_model(model, d)) for d in draws]
    say("config5-test: batch of %d draws" % len(draws))
    pl = LikelihoodPipeline(model, max_batch=len(draws))
    got = pl.evaluate(draws)
    pl.close()
    for k, (g, r) in enumerate(zip(got, ref)):
        say("  draw %d: gpu %.10g oracle %.10g" % (k, g, r))
        assert (g <= 0.1 * LLZERO) == (r <= 0.1 * LLZERO)
        if r > 0.1 * LLZERO:
            assert abs(g - r) < LNL_ATOL


def test_elsewhere_precision_form():
    from xpsi_amd import configs
    from xpsi_amd.pipeline import LikelihoodPipeline, params_from_model
    model = configs.config5("test")
    p = [params_from_model(model)]
    l64 = LikelihoodPipeline(dict(model, precision="fp64"), max_batch=1).evaluate(p)[0]
    l32 = LikelihoodPipeline(dict(model, precision="mixed"), max_batch=1).evaluate(p)[0]
    say("config5-test: lnL fp64 %.10f mixed %.10f |dlnL| %.3e" % (l64, l32, abs(l32 - l64)))
    assert abs(l32 - l64) < 1.0e-3


def test_time_invariant_dropin():
    import math
    from xpsi_amd.cellmesh.integrator_for_time_invariance import integrate
    st = O.spacetime(1.6, 11.0, 1.0, 0.3, 400.0)
    em, err = O.elsewhere_embed(st, dict(sqrt_num_cells=32, num_rays=500))
    assert err is None
    E = np.logspace(np.log10(0.3), 1.0, 96)
    t, ref = O.elsewhere_integrate(st, em, E, 5.9)
    assert t == 0
    S = em["S"]
    theta = np.repeat(em["colat"][:, None], S, axis=1)
    phi = np.repeat(em["phis"][None, :], S, axis=0)
    params = np.ones((S, S, 2))
    params[..., 0] = 5.9
    params[..., 1] = em["effGrav"][:, None]
    say("time-invariant drop-in: S=%d N_R=500 N_E=96" % S)
    status, flux = integrate(1, st["R"], st["Omega"], st["r_s"], st["i"], S, em["cellArea"], em["r"],
                             em["r_s_over_r"], theta, phi, params, 500, em["deflection"],
                             em["cos_alpha"], em["maxDeflection"], em["cos_gamma"], E, (), 1)
    assert status == 0
    e = relerr(flux, ref)
    say("time-invariant drop-in rel err %.3e" % e)
    assert e < SIG_RTOL


def test_integrate_dropin_with_correction():
    import math
    from xpsi_amd import configs
    from xpsi_amd.cellmesh.integrator_for_azimuthal_invariance import integrate
    model = configs.config1()
    sp = model["spacetime"]
    st = O.spacetime(sp["mass"], sp["radius"], sp["distance"], sp["cos_inclination"], sp["frequency"])
    hr = dict(model["hot_regions"][0], sqrt_num_cells=24)
    members, err = O.hot_region_embed(st, hr, sp["frequency"])
    m = members[0]
    E = O.construct_energy_array(64, model["instrument"]["energy_edges"])
    leaves = np.linspace(0.0, 2.0 * math.pi, 64)
    phases = 2.0 * math.pi * np.linspace(0.0, 1.0, 64)
    corr = np.ones_like(m["params"])
    corr[..., 0] = 5.8
    corr[..., 1] = m["effGrav"][:, None]
    t, ref = O.integrate(st["Omega"], st["r_s"], st["i"], m["cellArea"], m["r"], m["r_s_over_r"],
                         m["theta"], m["phi"], m["params"], m["radiates"], m["deflection"],
                         m["cos_alpha"], m["lag"], m["maxDeflection"], m["cos_gamma"], E, leaves,
                         phases, correction=corr)
    assert t == 0
    for kind in (1, 2):
        from xpsi_amd._lib import lib
        lib().xh_set_fold_kernel(kind)
        say("integrate drop-in with correction, fold kernel %d" % kind)
        status, flux = integrate(1, st["R"], st["Omega"], st["r_s"], st["i"], m["cellArea"], m["r"],
                                 m["r_s_over_r"], m["theta"], m["phi"], m["params"], m["radiates"],
                                 corr, m["deflection"].shape[1], m["deflection"], m["cos_alpha"],
                                 m["lag"], m["maxDeflection"], m["cos_gamma"], E, leaves, phases,
                                 (), (), 1, 1, 0)
        lib().xh_set_fold_kernel(0)
        assert status == 0
        e = relerr(flux, ref)
        say("integrate drop-in with correction rel err %.3e" % e)
        assert e < SIG_RTOL
