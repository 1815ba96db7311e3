"""GPU precision form against the FP64 path, and the two FP64 fold kernels
against each other.

The precision form computes the blackbody intensity, the spline
coefficients, the interval moments and their products in FP32 (matrix
cores), folds the response on FP32 matrix cores with FP64 sums of each
4-deep slice, and keeps the ray tables, the per-parallel phase-bin
accumulation and the likelihood quadrature in FP64.  Its bar is
|lnL - lnL_FP64| < 1e-3, asserted for the default response fold and for
the six-product split-BF16 one; the three-product split-BF16 fold is
printed as a measurement only (its ~5e-6 signal error exceeds the bar on
badly fitting draws).
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests.batch_cases import sample

pytestmark = pytest.mark.gpu

PREC_ATOL = 1.0e-3
LNL_ATOL = 1.0e-8
LLZERO = -1.0e90


def say(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _run(model, params, precision, fold_kernel="auto", fold_terms=0):
    from xpsi_amd.pipeline import LikelihoodPipeline
    m = dict(model, precision=precision, fold_kernel=fold_kernel, fold_terms=fold_terms)
    pl = LikelihoodPipeline(m, max_batch=len(params))
    lnl = pl.evaluate(params)
    nreg = len(model["hot_regions"])
    folded = [[pl.fetch("folded", k, r, pl.N_ch * pl.N_P)[:pl.N_ch * pl.N_P] for r in range(nreg)]
              for k in range(len(params))]
    pl.close()
    return np.asarray(lnl), folded


def _single(model, name):
    from xpsi_amd.pipeline import params_from_model
    lnl_ref = O.loglikelihood(model)
    p = [params_from_model(model)]
    say("%s: FP64 automatic fold" % name)
    l64, f64 = _run(model, p, "fp64")
    say("%s: FP64 per-point fold" % name)
    lpp, fpp = _run(model, p, "fp64", "per-point")
    say("%s: FP64 moment fold" % name)
    lmo, fmo = _run(model, p, "fp64", "moment")
    say("%s: precision form" % name)
    l32, f32 = _run(model, p, "mixed")
    l6, _ = _run(model, p, "mixed", fold_terms=6)
    l3, f3 = _run(model, p, "mixed", fold_terms=3)
    for r in range(len(f64[0])):
        e_pp = max(relerr(f64[0][r], fpp[0][r]), relerr(fmo[0][r], fpp[0][r]))
        e_32 = relerr(f32[0][r], f64[0][r])
        e_3 = relerr(f3[0][r], f64[0][r])
        say("%s region %d: folded rel err automatic/moment-vs-per-point %.2e, precision-vs-FP64 %.2e "
            "(3-product response fold %.2e)" % (name, r, e_pp, e_32, e_3))
        assert e_pp < 1.0e-10
    say("%s: lnL oracle %.10f fp64 %.10f per-point %.10f mixed %.10f |dlnL| %.3e "
        "(bf16 response fold: 6 products %.3e, 3 products %.3e)"
        % (name, lnl_ref, l64[0], lpp[0], l32[0], abs(l32[0] - l64[0]), abs(l6[0] - l64[0]),
           abs(l3[0] - l64[0])))
    assert abs(l64[0] - lnl_ref) < LNL_ATOL
    assert abs(lpp[0] - lnl_ref) < LNL_ATOL
    assert abs(lmo[0] - lnl_ref) < LNL_ATOL
    assert abs(l32[0] - l64[0]) < PREC_ATOL
    assert abs(l6[0] - l64[0]) < PREC_ATOL


def test_precision_config1():
    from xpsi_amd import configs
    _single(configs.config1(), "config1")


def test_precision_config2_bb():
    from xpsi_amd import configs
    _single(configs.config2("BB"), "config2-BB")


def test_precision_config2_num4d():
    from xpsi_amd import configs
    _single(configs.config2("Num4D"), "config2-Num4D")


def test_precision_batched_sample():
    from xpsi_amd import configs
    cs = sample(24, max_cells=96, max_e=256)
    groups = {}
    for k, d in enumerate(cs):
        groups.setdefault(d["num_energies"], []).append(k)
    worst = worst6 = worst3 = 0.0
    for ne, ks in sorted(groups.items()):
        model = configs.config3_model(ne)
        ps = [cs[k] for k in ks]
        say("batch of %d draws, %d energies" % (len(ks), ne))
        l64, _ = _run(model, ps, "fp64")
        lpp, _ = _run(model, ps, "fp64", "per-point")
        l32, _ = _run(model, ps, "mixed")
        l6, _ = _run(model, ps, "mixed", fold_terms=6)
        l3, _ = _run(model, ps, "mixed", fold_terms=3)
        for k, a, b, c, c6, c3 in zip(ks, l64, lpp, l32, l6, l3):
            say("  draw %d: fp64 %.10g per-point %.10g mixed %.10g (bf16 6-product %.10g, 3-product %.10g)"
                % (k, a, b, c, c6, c3))
            assert (a <= 0.1 * LLZERO) == (b <= 0.1 * LLZERO) == (c <= 0.1 * LLZERO)
            if a > 0.1 * LLZERO:
                assert abs(a - b) < LNL_ATOL
                worst = max(worst, abs(c - a))
                worst6 = max(worst6, abs(c6 - a))
                worst3 = max(worst3, abs(c3 - a))
                assert abs(c - a) < PREC_ATOL
                assert abs(c6 - a) < PREC_ATOL
    say("batched sample: max |dlnL| precision vs FP64 %.3e (bf16 response fold: 6 products %.3e, "
        "3 products %.3e)" % (worst, worst6, worst3))


def test_precision_ragged_tiles():
    """40 leaves (the 32-phase tiles of the spline and response folds end
    part-filled), 90 channels and 126 input bins (the response folds' 32-row
    tiles and 4- and 32-deep K steps end part-filled): FP64 forms against the
    oracle and each other, every precision-form response fold against FP64."""
    from xpsi_amd import configs
    base = configs.config1()
    M, edges, _ = configs.gaussian_response(90, 126)
    model = dict(base, hot_regions=[dict(base["hot_regions"][0], num_leaves=40)],
                 instrument=dict(matrix=M, energy_edges=edges),
                 data=dict(base["data"], counts=np.ascontiguousarray(base["data"]["counts"][:90])))
    _single(model, "ragged (40 leaves, 90 x 126 response)")
