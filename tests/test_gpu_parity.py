"""GPU parity: the HIP path against the CPU oracle on the same inputs.

Bars (from the path's specification): signal arrays agree to 1e-10 relative,
lnL to 1e-8 absolute, FP64 throughout.  Arrays are compared element by
element (tests/bars.py: each element relative to its own magnitude, floored
at 1e-6 of the array's largest).
"""
import math
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests.bars import elem_relerr, norm_relerr
from tests.models import examples_fast_model, examples_fast_support

pytestmark = pytest.mark.gpu

SIG_RTOL = 1.0e-10
LNL_ATOL = 1.0e-8


def say(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


relerr = elem_relerr


def oracle_member(model):
    sp = model["spacetime"]
    st = O.spacetime(sp["mass"], sp["radius"], sp["distance"], sp["cos_inclination"],
                     sp["frequency"], sp.get("star_shape", "AGM_14"))
    hr = model["hot_regions"][0]
    members, err = O.hot_region_embed(st, hr, sp["frequency"])
    assert err is None
    return st, hr, members


@pytest.fixture(scope="module")
def ef_model(examples_fast):
    return examples_fast_model(examples_fast)


@pytest.fixture(params=["per-point", "moment", "lanes"])
def fold_kernel(request):
    from xpsi_amd._lib import lib
    kind = {"per-point": 1, "moment": 2, "lanes": 3}[request.param]
    assert lib().xh_set_fold_kernel(kind) == 0
    yield request.param
    lib().xh_set_fold_kernel(0)


def test_integrate_dropin(ef_model, fold_kernel):
    from xpsi_amd.cellmesh.integrator_for_azimuthal_invariance import integrate
    st, hr, members = oracle_member(ef_model)
    m = members[0]
    energies = O.construct_energy_array(128, ef_model["instrument"]["energy_edges"])
    leaves = np.linspace(0.0, 2.0 * math.pi, 64)
    phases = 2.0 * math.pi * np.linspace(0.0, 1.0, 64)
    t, ref = O.integrate(st["Omega"], st["r_s"], st["i"], m["cellArea"], m["r"], m["r_s_over_r"],
                         m["theta"], m["phi"], m["params"], m["radiates"], m["deflection"],
                         m["cos_alpha"], m["lag"], m["maxDeflection"], m["cos_gamma"], energies,
                         leaves, phases, image_order_limit=3)
    assert t == 0
    say("integrate drop-in (%s fold): S=%d N_R=%d N_E=128 N_L=64"
        % (fold_kernel, m["S"], m["deflection"].shape[1]))
    status, flux = integrate(1, st["R"], st["Omega"], st["r_s"], st["i"], m["cellArea"], m["r"],
                             m["r_s_over_r"], m["theta"], m["phi"], m["params"], m["radiates"],
                             None, m["deflection"].shape[1], m["deflection"], m["cos_alpha"],
                             m["lag"], m["maxDeflection"], m["cos_gamma"], energies, leaves,
                             phases, (), (), 1, None, 0, image_order_limit=3)
    assert status == 0
    err = relerr(flux, ref)
    say("integrate drop-in rel err %.3e" % err)
    assert err < SIG_RTOL


def test_energy_integrator_dropin():
    from xpsi_amd.tools import energy_integrator
    rng = np.random.default_rng(7)
    E = np.logspace(np.log10(0.3), np.log10(10.0), 96)
    sig = np.exp(-((np.log10(E)[:, None] - 0.1) ** 2) / 0.3) * (1.0 + 0.3 * rng.random((96, 20)))
    sig[:, 3] = 0.0                       # a hidden column exercises the flat branch
    edges = np.logspace(np.log10(0.3), np.log10(12.0), 50)
    ref = O.energy_integrator(sig, np.log10(E), np.log10(edges))
    say("energy integrator drop-in: 96 energies x 20 phases -> 49 bins")
    out = energy_integrator(1, sig, np.log10(E), np.log10(edges))
    err = relerr(out, ref)
    say("energy integrator rel err %.3e" % err)
    assert err < SIG_RTOL
    assert np.all(out[np.log10(edges[:-1]) > np.log10(E[-1])] == 0.0)


def test_marginal_dropin(examples_fast):
    from xpsi_amd.likelihoods import default_background_marginalisation as G
    model = examples_fast_model(examples_fast)
    lnl_ref, parts = O.loglikelihood(model, return_parts=True)
    sig = parts["signals"]
    counts = model["data"]["counts"]
    pre = O.precomputation(counts.astype(np.int32))
    pre_g = G.precomputation(counts.astype(np.int32))
    assert np.allclose(pre_g, pre, rtol=1e-14, atol=1e-9)
    sup = -np.ones((counts.shape[0], 2)); sup[:, 0] = 0.0
    cph = [np.linspace(0.0, 1.0, 64)]
    say("marginal drop-in: 291 channels x 32 bins, unbounded support")
    out = G.eval_marginal_likelihood(1000.0, np.linspace(0, 1, 33), counts, sig, cph,
                                     np.array([0.0]), pre, sup, 1000, 0.0, 1.0e-8, 1.0e-3,
                                     10.0, -1.0e90)
    ref = O.eval_marginal_likelihood(1000.0, np.linspace(0, 1, 33), counts, sig, cph,
                                     np.array([0.0]), pre, sup)
    say("marginal lnL gpu %.12f oracle %.12f" % (out[0], ref[0]))
    assert abs(out[0] - ref[0]) < LNL_ATOL
    assert relerr(out[1], ref[1]) < 1e-9
    assert relerr(out[2], ref[2]) < 1e-9
    # bounded support (the examples_fast main.py recipe)
    sup2 = examples_fast_support(examples_fast["background_counts"])
    say("marginal drop-in: bounded support")
    out2 = G.eval_marginal_likelihood(1000.0, np.linspace(0, 1, 33), counts, sig, cph,
                                      np.array([0.0]), pre, sup2, 1000, 0.0, 1.0e-8, 1.0e-3,
                                      10.0, -1.0e90)
    ref2 = O.eval_marginal_likelihood(1000.0, np.linspace(0, 1, 33), counts, sig, cph,
                                      np.array([0.0]), pre, sup2)
    assert abs(out2[0] - ref2[0]) < LNL_ATOL


def _pipeline_case(model, name):
    from xpsi_amd.pipeline import LikelihoodPipeline, params_from_model
    lnl_ref, parts = O.loglikelihood(model, return_parts=True)
    st, hr, members = oracle_member(model)
    say("pipeline %s: creating" % name)
    pl = LikelihoodPipeline(model, max_batch=1)
    say("pipeline %s: evaluating" % name)
    lnl = pl.evaluate([params_from_model(model)])[0]
    say("pipeline %s: lnL gpu %.10f oracle %.10f" % (name, lnl, lnl_ref))
    m = members[0]
    S = m["S"]
    sizes = pl.fetch("sizes", 0, 0, 3)
    assert int(sizes[0]) == S and int(sizes[1]) == 0
    ca = pl.fetch("cellArea", 0, 0, S * S)[:S * S].reshape(S, S)
    assert relerr(ca, m["cellArea"]) < SIG_RTOL
    th = pl.fetch("theta", 0, 0, S)[:S]
    assert relerr(th, m["theta"][:, 0]) < 1e-12
    N_R = m["deflection"].shape[1]
    de = pl.fetch("deflection", 0, 0, S * N_R)[:S * N_R].reshape(S, N_R)
    assert relerr(de, m["deflection"]) < SIG_RTOL
    lg = pl.fetch("lag", 0, 0, S * N_R)[:S * N_R].reshape(S, N_R)
    assert relerr(lg, m["lag"]) < SIG_RTOL
    fo = pl.fetch("folded", 0, 0, pl.N_ch * pl.N_P)[:pl.N_ch * pl.N_P].reshape(pl.N_ch, pl.N_P)
    err = relerr(fo, parts["signals"][0])
    say("pipeline %s: folded signal rel err %.3e" % (name, err))
    assert err < SIG_RTOL
    assert abs(lnl - lnl_ref) < LNL_ATOL
    pl.close()
    return lnl


def test_pipeline_examples_fast(ef_model):
    lnl = _pipeline_case(ef_model, "examples_fast")
    assert np.isclose(lnl, -47881.27817666349, rtol=1.0e-5)


def test_pipeline_config1():
    from xpsi_amd import configs
    _pipeline_case(configs.config1(), "config1")


def _region_parity(model, name, lnl_tol=LNL_ATOL):
    """Full pipeline with every region and member checked against the oracle."""
    from xpsi_amd.pipeline import LikelihoodPipeline, params_from_model
    lnl_ref, parts = O.loglikelihood(model, return_parts=True)
    say("pipeline %s: evaluating (%d regions)" % (name, len(model["hot_regions"])))
    pl = LikelihoodPipeline(model, max_batch=1)
    lnl = pl.evaluate([params_from_model(model)])[0]
    say("pipeline %s: lnL gpu %.10f oracle %.10f" % (name, lnl, lnl_ref))
    for r in range(len(model["hot_regions"])):
        fo = pl.fetch("folded", 0, r, pl.N_ch * pl.N_P)[:pl.N_ch * pl.N_P].reshape(pl.N_ch, pl.N_P)
        err = relerr(fo, parts["signals"][r])
        say("pipeline %s: region %d folded rel err %.3e" % (name, r, err))
        assert err < SIG_RTOL
    assert abs(lnl - lnl_ref) < lnl_tol
    pl.close()


def test_pipeline_config2_bb_two_regions_cede():
    from xpsi_amd import configs
    _region_parity(configs.config2("BB"), "config2-BB")


def test_pipeline_config2_num4d():
    from xpsi_amd import configs
    _region_parity(configs.config2("Num4D"), "config2-Num4D")


def test_integrate_dropin_num4d():
    from xpsi_amd import configs
    from xpsi_amd.cellmesh.integrator_for_azimuthal_invariance import integrate
    model = configs.config2("Num4D")
    sp = model["spacetime"]
    st = O.spacetime(sp["mass"], sp["radius"], sp["distance"], sp["cos_inclination"],
                     sp["frequency"])
    hr = model["hot_regions"][0]
    members, err = O.hot_region_embed(st, hr, sp["frequency"])
    atm = model["atmosphere"]
    energies = O.construct_energy_array(256, model["instrument"]["energy_edges"])
    leaves = np.linspace(0.0, 2.0 * math.pi, 128)
    phases = 2.0 * math.pi * np.linspace(0.0, 1.0, 128)
    for k, m in enumerate(members):
        t, ref = O.integrate(st["Omega"], st["r_s"], st["i"], m["cellArea"], m["r"],
                             m["r_s_over_r"], m["theta"], m["phi"], m["params"], m["radiates"],
                             m["deflection"], m["cos_alpha"], m["lag"], m["maxDeflection"],
                             m["cos_gamma"], energies, leaves, phases, atm_ext=2, atmosphere=atm)
        assert t == 0
        say("integrate drop-in Num4D member %d: S=%d" % (k, m["S"]))
        status, flux = integrate(1, st["R"], st["Omega"], st["r_s"], st["i"], m["cellArea"], m["r"],
                                 m["r_s_over_r"], m["theta"], m["phi"], m["params"],
                                 m["radiates"], None, m["deflection"].shape[1], m["deflection"],
                                 m["cos_alpha"], m["lag"], m["maxDeflection"], m["cos_gamma"],
                                 energies, leaves, phases, atm, (), 2, None, 0)
        assert status == 0
        err = relerr(flux, ref)
        say("integrate drop-in Num4D member %d rel err %.3e" % (k, err))
        assert err < SIG_RTOL


def test_pipeline_omitted_region():
    # HotRegion omit=True (xpsi/HotRegion.py:30-48, mesh_tools.pyx:892-910): a
    # circular region cut out of the superseding spot, the cells' azimuths
    # shifted by -omit_azimuth; checked against the oracle, and the hole must
    # take flux away
    from xpsi_amd import configs
    m = configs.config1()
    hr = dict(m["hot_regions"][0], omit=True, omit_colatitude=0.85, omit_radius=0.15, omit_azimuth=0.1)
    model = dict(m, hot_regions=[hr])
    _region_parity(model, "config1 with an omitted region")
    from xpsi_amd.pipeline import LikelihoodPipeline, params_from_model
    pl = LikelihoodPipeline(m, max_batch=1)
    pl.evaluate([params_from_model(m)])
    full = pl.fetch("folded", 0, 0, pl.N_ch * pl.N_P)[:pl.N_ch * pl.N_P]
    pl.close()
    pl = LikelihoodPipeline(model, max_batch=1)
    pl.evaluate([params_from_model(model)])
    cut = pl.fetch("folded", 0, 0, pl.N_ch * pl.N_P)[:pl.N_ch * pl.N_P]
    pl.close()
    say("omitted region: folded signal %.4f of the full spot's" % (cut.sum() / full.sum()))
    assert 0.5 < cut.sum() / full.sum() < 0.95
