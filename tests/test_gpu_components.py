"""GPU: the likelihood drop-ins with components on different phase grids and
allow_negative per component (xpsi/tools/compute_expected_counts.pyx:26-197,
xpsi/likelihoods/default_background_marginalisation.pyx:450-734,
xpsi/likelihoods/_poisson_likelihood_given_background.pyx:14-113) against the
oracle, which takes both as the reference does.  Two components: one positive
on 32 phases, one with negative stretches on 47 phases (a different grid),
shifted; each allow_negative pattern and the bool forms; lnL to 1e-8
absolute, the expected-count arrays elementwise to 1e-10."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests.bars import LNL_ATOL, SIG_RTOL, elem_relerr

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


def _case(seed=5, N_ch=24, N_b=16):
    rng = np.random.default_rng(seed)
    g0 = np.linspace(0.0, 1.0, 32)
    g1 = np.sort(np.concatenate([[0.0, 1.0], rng.uniform(0.02, 0.98, 45)]))
    amp = rng.uniform(0.5, 2.0, N_ch)[:, None]
    c0 = amp * (1.0 + 0.6 * np.cos(2 * np.pi * g0))[None, :]
    c1 = amp * (0.35 * np.sin(2 * np.pi * g1) - 0.1)[None, :]
    c1[:, -1] = c1[:, 0]          # periodic grids repeat their first value
    bg = rng.uniform(0.2, 0.6, N_ch)[:, None] * np.ones(N_b)
    phases = np.linspace(0.0, 1.0, N_b + 1)
    shifts = np.array([0.07, -0.31])
    lam = 100.0 * (amp * (1.3 + 0.2 * np.cos(2 * np.pi * phases[:-1]))[None, :]) + 100.0 * bg / N_b
    counts = rng.poisson(lam).astype(np.float64)
    return phases, counts, [c0, c1], [g0, g1], shifts, bg


@pytest.mark.parametrize("allow", [False, True, [True, False], [False, True]])
def test_poisson_components_on_their_own_grids(allow):
    from xpsi_amd.likelihoods._poisson_likelihood_given_background import poisson_likelihood_given_background as G
    phases, counts, comps, grids, shifts, bg = _case()
    say("poisson, grids 32 and 47 phases, allow_negative %s" % (allow,))
    ref = O.poisson_likelihood_given_background(100.0, phases, counts, comps, grids, shifts, bg,
                                                allow_negative=allow)
    got = G(100.0, phases, counts, comps, grids, shifts, bg, allow_negative=allow)
    say("lnL gpu %.10f oracle %.10f" % (got[0], ref[0]))
    assert abs(got[0] - ref[0]) < LNL_ATOL
    assert elem_relerr(got[1], ref[1]) < SIG_RTOL


@pytest.mark.parametrize("allow", [False, [True, False], [False, True]])
def test_marginal_components_on_their_own_grids(allow):
    from xpsi_amd.likelihoods import default_background_marginalisation as G
    phases, counts, comps, grids, shifts, bg = _case(seed=7)
    pre = O.precomputation(counts.astype(np.int32))
    sup = np.zeros((counts.shape[0], 2))
    sup[:, 1] = -1.0
    say("marginal, grids 32 and 47 phases, allow_negative %s" % (allow,))
    out = G.eval_marginal_likelihood(100.0, phases, counts, comps, grids, shifts, pre, sup, 1000, 0.0, 1.0e-8,
                                     1.0e-3, 10.0, -1.0e90, allow_negative=allow)
    ref = O.eval_marginal_likelihood(100.0, phases, counts, comps, grids, shifts, pre, sup,
                                     allow_negative=allow)
    say("lnL gpu %.10f oracle %.10f" % (out[0], ref[0]))
    assert abs(out[0] - ref[0]) < LNL_ATOL
    assert elem_relerr(out[1], ref[1]) < SIG_RTOL


def test_allow_negative_count_mismatch_raises():
    from xpsi_amd.likelihoods._poisson_likelihood_given_background import poisson_likelihood_given_background as G
    phases, counts, comps, grids, shifts, bg = _case()
    with pytest.raises(ValueError):
        G(100.0, phases, counts, comps, grids, shifts, bg, allow_negative=[True])
