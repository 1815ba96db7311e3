"""GPU: prior inverse sampling on the device (xpsi/Prior.py:84-160 with the
examples' compactness prior, CustomPrior.py:49-66), the reference-shaped
Likelihood call with its change detection (Likelihood.py:450-513), and prior
draws fed into the continuous-batching server."""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOUNDS = [(1.0, 3.0), (8.0, 16.0), (1.0, 1.0), (0.0, 1.0), (0.0, 1.0), (0.0, np.pi), (0.05, 1.0),
          (5.5, 6.7)]


def say(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


def test_inverse_sample_and_compactness():
    from xpsi_amd import configs
    from xpsi_amd.prior import DevicePrior
    from xpsi_amd.sampling import parameter_names
    names = parameter_names(configs.config1())
    pr = DevicePrior(BOUNDS, names, 300.0)
    rng = np.random.default_rng(12)
    u = rng.random((5000, len(names)))
    P, lp = pr.inverse_sample(u)
    b = np.array(BOUNDS)
    np.testing.assert_allclose(P, b[:, 0] + (b[:, 1] - b[:, 0]) * u, rtol=1e-15, atol=0.0)
    acc = np.array([configs.prior_accepts(m, r, 300.0) for m, r in zip(P[:, 0], P[:, 1])])
    say("device prior: %d of 5000 accepted" % int(np.isfinite(lp).sum()))
    assert np.array_equal(np.isfinite(lp), acc)
    assert np.all(lp[acc] == 0.0) and np.all(np.isneginf(lp[~acc]))
    d = pr.draw(300, rng)
    assert d.shape == (300, len(names))
    assert all(configs.prior_accepts(m, r, 300.0) for m, r in zip(d[:, 0], d[:, 1]))


def test_likelihood_call_change_detection():
    from xpsi_amd import configs
    from xpsi_amd.pipeline import LikelihoodPipeline, params_from_model
    from xpsi_amd.sampling import Likelihood, parameter_names, vector_to_params
    m = configs.config1()
    names = parameter_names(m)
    p0 = params_from_model(m)
    h = p0["hot_regions"][0]
    v0 = np.array([p0[k] for k in names[:4]] + [h[k] for k in names[4:]])
    L = Likelihood(m)
    a = L(v0)
    b = L(v0)
    c = L(v0, force=True)
    pl = LikelihoodPipeline(m, max_batch=1)
    ref = pl.evaluate([vector_to_params(m, v0)])[0]
    say("Likelihood: %.10f (again %.10f, forced %.10f), pipeline %.10f" % (a, b, c, ref))
    assert a == b and abs(c - a) < 1e-9 and abs(a - ref) < 1e-9
    bad = v0.copy()
    bad[2], bad[6], bad[7] = 0.05, 1.0, 6.9        # near, large and hot: llzero-class
    x = L(bad)
    y = L(bad)
    z = L(bad, force=True)
    assert -1e90 <= x <= -1e89 and x == y and z != x and -1e90 <= z <= -1e89
    assert L() == L(force=True) or L() <= -1e89    # p=None re-evaluates the current vector
    got = L.batch(np.stack([v0, bad, v0]))
    assert abs(got[0] - ref) < 1e-9 and abs(got[2] - ref) < 1e-9 and got[1] <= -1e89
    pl.close()
    L.close()


def test_prior_draws_fed_into_server():
    from xpsi_amd import configs
    from xpsi_amd.pipeline import LikelihoodPipeline
    from xpsi_amd.prior import DevicePrior
    from xpsi_amd.sampling import parameter_names, serve_prior_draws, vector_to_params
    from xpsi_amd.serve import LikelihoodServer
    m = configs.config1()
    names = parameter_names(m)
    pr = DevicePrior(BOUNDS, names, 300.0)
    srv = LikelihoodServer(m, n_slots=2, slot_batch=16, use_graphs=True, max_wait_s=0.001)
    vecs, ids = serve_prior_draws(srv, m, pr, 40, np.random.default_rng(3), batch=64)
    got = {}
    t0 = time.time()
    while len(got) < len(ids) and time.time() - t0 < 60:
        i, l, _ = srv.poll()
        got.update(zip(i.tolist(), l.tolist()))
        time.sleep(0.001)
    srv.close()
    assert len(got) == len(ids)
    res = np.array([got[int(i)] for i in ids])
    pl = LikelihoodPipeline(m, max_batch=1)
    ref = np.array([pl.evaluate([vector_to_params(m, v)])[0] for v in vecs])
    pl.close()
    fin = ref > -1e89
    say("server fed by device prior draws: %d of %d finite" % (int(fin.sum()), len(ref)))
    assert np.array_equal(fin, res > -1e89)
    assert np.max(np.abs(res[fin] - ref[fin])) < 1e-9
