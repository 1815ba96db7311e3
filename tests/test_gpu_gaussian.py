"""GPU: the Gaussian likelihoods against the oracle.

gaussian_likelihood_given_background
(xpsi/likelihoods/_gaussian_likelihood_given_background_IQU.pyx:14-91): lnL to
1e-8 absolute and 1e-12 relative, expected counts element by element to 1e-10,
on the reference's own test model's signals (two shifted components), with
allow_negative per component and with the components on different phase
grids.  gaussian_likelihood_QnUn (_gaussian_likelihood_QnUn.pyx:14-44): lnL
to 1e-12 relative on Stokes-like rows about zero."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests.bars import LNL_ATOL, SIG_RTOL, elem_relerr
from tests.models import examples_fast_model

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


def _case(examples_fast):
    model = examples_fast_model(examples_fast)
    _, parts = O.loglikelihood(model, return_parts=True)
    sig = parts["signals"][0]
    counts = np.ascontiguousarray(model["data"]["counts"], dtype=np.float64)
    N_ch, N_b = counts.shape
    phases = np.linspace(0.0, 1.0, N_b + 1)
    cph = np.linspace(0.0, 1.0, sig.shape[1])
    bg = np.ascontiguousarray(examples_fast["background_counts"] / 1000.0 / N_b)[:, None] * np.ones(N_b)
    errors = np.sqrt(np.maximum(counts, 1.0)) * (1.0 + 0.1 * np.cos(np.arange(N_b)))[None, :]
    return sig, counts, phases, cph, bg, errors


def _check(got, ref, what):
    say("%s: lnL gpu %.10f oracle %.10f" % (what, got[0], ref[0]))
    assert abs(got[0] - ref[0]) < LNL_ATOL
    assert abs(got[0] - ref[0]) <= 1.0e-12 * abs(ref[0])
    assert elem_relerr(got[1], ref[1]) < SIG_RTOL


def test_gaussian_given_background_vs_oracle(examples_fast):
    from xpsi_amd.likelihoods._gaussian_likelihood_given_background_IQU import (
        gaussian_likelihood_given_background as gpu_gl)
    sig, counts, phases, cph, bg, errors = _case(examples_fast)
    comps = [sig, 0.3 * np.roll(sig, 7, axis=1)]
    shifts = np.array([0.05, -0.3])
    say("gaussian given background: %d channels x %d bins" % counts.shape)
    ref = O.gaussian_likelihood_given_background(1000.0, phases, counts, errors, comps, [cph, cph], shifts, bg)
    got = gpu_gl(1000.0, phases, counts, errors, comps, [cph, cph], shifts, bg)
    _check(got, ref, "shared grid")
    # the second component negated: dropped without its flag, kept with it
    neg = [sig, -0.3 * np.roll(sig, 7, axis=1)]
    for flags in ([False, False], [False, True]):
        ref = O.gaussian_likelihood_given_background(1000.0, phases, counts, errors, neg, [cph, cph], shifts, bg,
                                                     allow_negative=flags)
        got = gpu_gl(1000.0, phases, counts, errors, neg, [cph, cph], shifts, bg, allow_negative=flags)
        _check(got, ref, "allow_negative %s" % flags)


def test_gaussian_given_background_component_grids(examples_fast):
    from xpsi_amd.likelihoods._gaussian_likelihood_given_background_IQU import (
        gaussian_likelihood_given_background as gpu_gl)
    sig, counts, phases, cph, bg, errors = _case(examples_fast)
    # a second component on a coarser grid of its own
    cph2 = np.linspace(0.0, 1.0, 17)
    second = np.stack([np.interp(cph2, cph, row) for row in sig]) * 0.5
    shifts = np.array([0.1, 0.45])
    ref = O.gaussian_likelihood_given_background(1000.0, phases, counts, errors, [sig, second], [cph, cph2],
                                                 shifts, bg)
    got = gpu_gl(1000.0, phases, counts, errors, [sig, second], [cph, cph2], shifts, bg)
    _check(got, ref, "own phase grids")


def test_gaussian_QnUn_vs_oracle():
    from xpsi_amd.likelihoods._gaussian_likelihood_QnUn import gaussian_likelihood_QnUn as gpu_q
    rng = np.random.default_rng(17)
    for rows, N_b in ((1, 32), (4, 100), (2, 1)):
        star = rng.uniform(-0.4, 0.4, N_b)
        counts = star[None, :] + rng.normal(0.0, 0.03, (rows, N_b))
        errors = rng.uniform(0.02, 0.05, (rows, N_b))
        edges = np.linspace(0.0, 1.0, N_b + 1)
        ref = O.gaussian_likelihood_QnUn(edges, counts, errors, star)
        got = gpu_q(edges, counts, errors, star)
        say("QnUn %d x %d: lnL gpu %.12f oracle %.12f" % (rows, N_b, got[0], ref[0]))
        assert abs(got[0] - ref[0]) <= 1.0e-12 * abs(ref[0])
        np.testing.assert_array_equal(got[1], star)
