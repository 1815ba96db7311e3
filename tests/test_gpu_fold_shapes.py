"""Every block shape of the moment-form fold and the energy-lane fold give the
same likelihoods.

The fold kernel's block shape (T phase tiles x T resident energy chunks, with
the register budget sized for two or four waves per SIMD by the LDS the shape
needs) is tuning state of the library.  FP64: each shape against the per-point
fold to the path's lnL bar (1e-8) and signals to 1e-10.  Precision form: each
shape and the energy-lane fold against FP64 to the precision bar (1e-3); they
differ from each other only by FP32 rounding order (measured ~1e-4), checked
against the same bar.
"""
import sys

import numpy as np
import pytest

from tests.batch_cases import sample

pytestmark = pytest.mark.gpu

LNL_ATOL = 1.0e-8
SIG_RTOL = 1.0e-10
PREC_ATOL = 1.0e-3
SHAPE_ATOL = 1.0e-3
LLZERO = -1.0e90


def say(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


def _run(model, params, precision, fold_kernel, shape=0):
    from xpsi_amd._lib import lib
    from xpsi_amd.pipeline import LikelihoodPipeline
    assert lib().xh_set_moment_shape(shape) == 0
    try:
        pl = LikelihoodPipeline(dict(model, precision=precision, fold_kernel=fold_kernel),
                                max_batch=len(params))
        lnl = np.asarray(pl.evaluate(params))
        folded = [pl.fetch("folded", k, 0, pl.N_ch * pl.N_P)[:pl.N_ch * pl.N_P] for k in range(len(params))]
        pl.close()
    finally:
        lib().xh_set_moment_shape(0)
    return lnl, folded


@pytest.fixture(scope="module")
def draws():
    from xpsi_amd import configs
    cs = [d for d in sample(16, max_cells=64, max_e=128) if d["num_energies"] == 128]
    assert len(cs) >= 3
    return configs.config3_model(128), cs


def test_fp64_shapes_against_per_point(draws):
    model, ps = draws
    say("FP64 per-point fold, %d draws" % len(ps))
    ref, fref = _run(model, ps, "fp64", "per-point")
    ok = ref > 0.1 * LLZERO
    for shape in (1, 2, 4):
        say("FP64 moment fold, shape %d" % shape)
        l, f = _run(model, ps, "fp64", "moment", shape)
        assert np.all((l > 0.1 * LLZERO) == ok)
        d = float(np.max(np.abs(l[ok] - ref[ok]))) if ok.any() else 0.0
        e = max(float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
                for a, b, o in zip(f, fref, ok) if o)
        say("  max |dlnL| %.3e, max signal rel err %.3e" % (d, e))
        assert d < LNL_ATOL
        assert e < SIG_RTOL


def test_precision_shapes_and_lanes(draws):
    model, ps = draws
    say("FP64 reference, %d draws" % len(ps))
    ref, _ = _run(model, ps, "fp64", "moment")
    ok = ref > 0.1 * LLZERO
    runs = {}
    for kernel, shape in (("moment", 1), ("moment", 2), ("moment", 4), ("lanes", 0)):
        say("precision form, %s fold, shape %d" % (kernel, shape))
        l, _ = _run(model, ps, "mixed", kernel, shape)
        assert np.all((l > 0.1 * LLZERO) == ok)
        d = float(np.max(np.abs(l[ok] - ref[ok]))) if ok.any() else 0.0
        say("  max |dlnL| vs FP64 %.3e" % d)
        assert d < PREC_ATOL
        runs[(kernel, shape)] = l
    base = runs[("moment", 1)]
    for key, l in runs.items():
        d = float(np.max(np.abs(l[ok] - base[ok]))) if ok.any() else 0.0
        say("  %s/%d vs moment/1: %.3e" % (key[0], key[1], d))
        assert d < SHAPE_ATOL
