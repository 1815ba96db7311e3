"""GPU: the vectorised sampler-facing likelihood (flat vectors, many at a
time) against one-at-a-time pipeline evaluations and the oracle."""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


def test_vectorized_likelihood_config1():
    from xpsi_amd import configs
    from xpsi_amd.pipeline import LikelihoodPipeline, params_from_model
    from xpsi_amd.sampling import VectorizedLikelihood, parameter_names, vector_to_params
    m = configs.config1()
    names = parameter_names(m)
    p0 = params_from_model(m)
    h = p0["hot_regions"][0]
    v0 = np.array([p0[k] for k in names[:4]] + [h[k] for k in names[4:]])
    rng = np.random.default_rng(5)
    step = np.array([0.05, 0.2, 0.0, 0.02, 0.05, 0.05, 0.02, 0.02])
    P = v0 + rng.normal(0.0, 1.0, (24, len(names))) * step
    P[5, 1] = 3.0    # a 3 km star: the compactness prior rejects it

    def log_prior(v):
        return 0.0 if configs.prior_accepts(v[0], v[1], 300.0) else -np.inf

    f = VectorizedLikelihood(m, max_batch=10, log_prior=log_prior, seed=1)
    got = f(P)
    pl = LikelihoodPipeline(m, max_batch=1)
    for k in range(len(P)):
        if not np.isfinite(log_prior(P[k])):
            say("vector %d: prior rejects, lnL %.4g" % (k, got[k]))
            assert f.llzero <= got[k] <= 0.1 * f.llzero
            continue
        ref = pl.evaluate([vector_to_params(m, P[k])])[0]
        say("vector %d: batched %.10f alone %.10f" % (k, got[k], ref))
        if ref <= 0.1 * f.llzero:     # random_near_llzero: the value itself is random
            assert f.llzero <= got[k] <= 0.1 * f.llzero
        else:
            assert abs(got[k] - ref) < 1.0e-9
    pl.close()
    l0 = f(v0)
    ref0 = O.loglikelihood(m)
    say("model vector: %.10f oracle %.10f" % (l0, ref0))
    assert isinstance(l0, float)
    assert abs(l0 - ref0) < 1.0e-8
    f.close()
