"""GPU: the wave-autonomous moment fold (k_wave_fold, fold_kernel "waves")
against the FP64 path and the block moment fold of the same precision form.

Both precision-form folds compute the same FP32 spline coefficients and
moments; they differ in summation order (one FP32 MFMA chain per interval
sweep here, two there) and in where the per-parallel sums meet (FP64
atomics here).  Bars: folded signals within 2e-6 of FP64 element by element
(above the path's 1e-6 floor; the precision form's own error is ~1e-7),
|lnL - lnL_FP64| < 1e-3, and the llzero class unchanged.  Cases: one region
(config 1), two regions with a ceding ring, ragged tiles (40 leaves and 40
phases, 100 energies: part-filled lanes on both sides), the elsewhere
correction (no clipping), and a batch of prior draws at mixed resolutions;
then the ragged, elsewhere and batch cases again with one block per parallel
whose energy chunks share the cell walk (k_wave_fold_shared), and that form
against the one-wavefront form."""
import sys

import numpy as np
import pytest

from tests.bars import elem_relerr
from tests.batch_cases import sample

pytestmark = pytest.mark.gpu

PREC_ATOL = 1.0e-3
SIG_BAR = 2.0e-6
LLZERO = -1.0e90


def say(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


def _run(model, params, precision, fold_kernel="auto"):
    from xpsi_amd.pipeline import LikelihoodPipeline
    m = dict(model, precision=precision, fold_kernel=fold_kernel)
    pl = LikelihoodPipeline(m, max_batch=len(params))
    lnl = pl.evaluate(params)
    nreg = len(model["hot_regions"])
    folded = [[pl.fetch("folded", k, r, pl.N_ch * pl.N_P)[:pl.N_ch * pl.N_P] for r in range(nreg)]
              for k in range(len(params))]
    pl.close()
    return np.asarray(lnl), folded


def _compare(model, params, name):
    say("%s: FP64" % name)
    l64, f64 = _run(model, params, "fp64")
    say("%s: precision form, block moment fold" % name)
    lmo, fmo = _run(model, params, "mixed", "moment")
    say("%s: precision form, wave fold" % name)
    lwv, fwv = _run(model, params, "mixed", "waves")
    worst_sig = worst_lnl = 0.0
    for k in range(len(params)):
        assert (l64[k] <= 0.1 * LLZERO) == (lwv[k] <= 0.1 * LLZERO), (k, l64[k], lwv[k])
        if l64[k] <= 0.1 * LLZERO:
            continue
        for r in range(len(f64[k])):
            e_wv = elem_relerr(fwv[k][r], f64[k][r])
            e_mo = elem_relerr(fmo[k][r], f64[k][r])
            worst_sig = max(worst_sig, e_wv)
            assert e_wv < SIG_BAR, (name, k, r, e_wv, e_mo)
        worst_lnl = max(worst_lnl, abs(lwv[k] - l64[k]))
        assert abs(lwv[k] - l64[k]) < PREC_ATOL, (name, k, lwv[k], l64[k], lmo[k])
    say("%s: wave fold vs FP64: folded %.2e (elementwise), |dlnL| %.3e; block fold |dlnL| %.3e"
        % (name, worst_sig, worst_lnl,
           max([abs(a - b) for a, b in zip(lmo, l64) if b > 0.1 * LLZERO] + [0.0])))


def test_wave_fold_config1():
    from xpsi_amd import configs
    from xpsi_amd.pipeline import params_from_model
    m = configs.config1()
    _compare(m, [params_from_model(m)], "config1")


def test_wave_fold_two_regions_cede():
    from xpsi_amd import configs
    from xpsi_amd.pipeline import params_from_model
    m = configs.config2("BB")
    _compare(m, [params_from_model(m)], "config2-BB")


def test_wave_fold_ragged():
    from xpsi_amd import configs
    cs = sample(6, max_cells=64, max_e=256)
    base = configs.config3_model(100)
    model = dict(base, hot_regions=[dict(base["hot_regions"][0], num_leaves=40, num_phases=40)])
    _compare(model, cs, "ragged (40 leaves, 40 phases, 100 energies)")


def test_wave_fold_elsewhere_correction():
    from xpsi_amd import configs
    m = configs.config5("test")
    draws = [d for d in configs.config5_draws(32) if configs.prior_accepts(d["mass"], d["radius"], 300.0)][:4]
    _compare(m, draws, "config5 test size (elsewhere correction)")


def test_wave_fold_batched_sample():
    from xpsi_amd import configs
    cs = sample(24, max_cells=96, max_e=256)
    groups = {}
    for k, d in enumerate(cs):
        groups.setdefault(d["num_energies"], []).append(k)
    for ne, ks in sorted(groups.items()):
        _compare(configs.config3_model(ne), [cs[k] for k in ks], "batch of %d draws, %d energies" % (len(ks), ne))


def _shared(fn):
    from xpsi_amd import lib
    from xpsi_amd._lib import check
    L = lib()
    check(L.xh_set_wave_fold_shared(1), "xh_set_wave_fold_shared")
    try:
        fn()
    finally:
        check(L.xh_set_wave_fold_shared(0), "xh_set_wave_fold_shared")


def test_wave_fold_shared_ragged():
    # 100 energies: two chunks, the second part-filled; 40 phases and leaves
    _shared(test_wave_fold_ragged)


def test_wave_fold_shared_elsewhere_correction():
    _shared(test_wave_fold_elsewhere_correction)


def test_wave_fold_shared_batched_sample():
    _shared(test_wave_fold_batched_sample)


def test_wave_fold_shared_matches_one_wavefront_form():
    """The two forms sum the same chains in the same order per energy: lnL
    within 1e-9 of each other on a batch at mixed resolutions."""
    from xpsi_amd import configs
    cs = sample(12, max_cells=96, max_e=256)
    ne = max(d["num_energies"] for d in cs)
    params = [d for d in cs if d["num_energies"] == ne]
    model = configs.config3_model(ne)
    say("shared walk vs one wavefront: %d draws, %d energies" % (len(params), ne))
    l1, _ = _run(model, params, "mixed", "waves")
    out = {}
    _shared(lambda: out.setdefault("lx", _run(model, params, "mixed", "waves")[0]))
    for a, b in zip(l1, out["lx"]):
        if a <= 0.1 * LLZERO:
            assert b <= 0.1 * LLZERO
            continue
        assert abs(a - b) < 1.0e-9 * max(1.0, abs(a)), (a, b)
