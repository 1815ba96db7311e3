"""GPU: the Poisson likelihood given a known background against the oracle
(xpsi/likelihoods/_poisson_likelihood_given_background.pyx:14-113): lnL to
1e-8 absolute, expected counts element by element to 1e-10, on the reference's
own test model's signals (two components, shifted), with and without the
factorial term, and the penalty of bins with data but no expected counts."""
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


def test_poisson_given_background_vs_oracle(examples_fast):
    from xpsi_amd.likelihoods._poisson_likelihood_given_background import (
        poisson_likelihood_given_background as gpu_pl)
    model = examples_fast_model(examples_fast)
    _, parts = O.loglikelihood(model, return_parts=True)
    sig = parts["signals"][0]
    counts = np.ascontiguousarray(model["data"]["counts"], dtype=np.float64)
    N_ch, N_b = counts.shape
    phases = np.linspace(0.0, 1.0, N_b + 1)
    cph = np.linspace(0.0, 1.0, sig.shape[1])
    comps = [sig, 0.3 * np.roll(sig, 7, axis=1)]
    shifts = np.array([0.05, -0.3])
    bg = np.ascontiguousarray(examples_fast["background_counts"] / 1000.0 / N_b)[:, None] * np.ones(N_b)
    pre = O.precomputation(counts.astype(np.int32))
    for use_pre in (True, False):
        p = pre if use_pre else None
        say("poisson given background: %d channels x %d bins, factorial term %s" % (N_ch, N_b, use_pre))
        ref = O.poisson_likelihood_given_background(1000.0, phases, counts, comps, [cph, cph], shifts,
                                                    bg, p)
        got = gpu_pl(1000.0, phases, counts, comps, [cph, cph], shifts, bg, p)
        say("lnL gpu %.10f oracle %.10f" % (got[0], ref[0]))
        assert abs(got[0] - ref[0]) < LNL_ATOL
        assert elem_relerr(got[1], ref[1]) < SIG_RTOL
    # data where nothing is expected: one draw near -1e90 per such bin, fresh per call
    c2 = counts.copy()
    z = [np.zeros_like(sig)]
    c2[:] = 0.0
    c2[3, 5] = 2.0
    a = gpu_pl(1000.0, phases, c2, z, [cph], shifts[:1], np.zeros_like(bg))[0]
    b = gpu_pl(1000.0, phases, c2, z, [cph], shifts[:1], np.zeros_like(bg))[0]
    say("penalised: %.6g %.6g" % (a, b))
    assert -1e90 <= a <= -1e89 and -1e90 <= b <= -1e89 and a != b
