"""The oracle reproduces the reference's own known-answer likelihoods.

xpsi/tests/test_likelihood.py:131-134 asserts -47881.27817666349 (rtol 1e-5);
examples/examples_fast/Modules/main.py:153-155 checks -3.2552616905e+04
(rtol 1e-5).  Both pin the whole path end to end: mesh, rays, ring
integrator, Akima/Steffen splines, energy integral, fold and marginal.
"""
import numpy as np

from oracle import oracle as O
from tests.models import examples_fast_model, examples_fast_support


def test_test_likelihood_known_answer(examples_fast):
    lnl = O.loglikelihood(examples_fast_model(examples_fast))
    assert np.isclose(lnl, -47881.27817666349, rtol=1.0e-5)
    # much tighter than the reference asks; a regression guard for the oracle
    assert abs(lnl + 47881.27817666349) < 1.0e-5


def test_examples_fast_main_known_answer(examples_fast):
    sup = examples_fast_support(examples_fast["background_counts"])
    lnl = O.loglikelihood(examples_fast_model(examples_fast, radius=12.0,
                                              logT=6.7, num_energies=64,
                                              support=sup))
    assert np.isclose(lnl, -3.2552616905e+04, rtol=1.0e-5)
