"""hipGraph replay and the continuous-batching server against eager,
one-at-a-time evaluation (FP64: identical within 1e-9)."""
import sys
import time

import numpy as np
import pytest

from tests.batch_cases import sample

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


def same_lnl(got, ref, tol):
    """Finite values agree within tol; llzero-class values (random draws in
    [llzero, 0.1 llzero], fresh per launch) agree as a class."""
    fin = ref > -1e89
    assert np.array_equal(fin, got > -1e89)
    assert np.all((got[~fin] >= -1e90) & (got[~fin] <= -1e89))
    if fin.any():
        assert np.max(np.abs(got[fin] - ref[fin])) < tol


def _draws(ne, n):
    from xpsi_amd import configs
    out = [d for d in configs.config3_draws()
           if configs.prior_accepts(d["mass"], d["radius"], 300.0) and d["num_energies"] == ne
           and d["hot_regions"][0]["sqrt_num_cells"] <= 64]
    return out[:n]


def test_graph_replay_matches_eager():
    import torch
    from xpsi_amd import configs
    from xpsi_amd.pipeline import LikelihoodPipeline
    draws = _draws(64, 12)
    model = configs.config3_model(64)
    eager = LikelihoodPipeline(model, max_batch=12)
    ref = eager.evaluate(draws)
    pl = LikelihoodPipeline(model, max_batch=12)
    pl.use_graphs(True)
    s = torch.cuda.Stream()
    for rep in range(3):
        say("graph replay %d (batch 12, then 5)" % rep)
        got = pl.evaluate(draws, s.cuda_stream)
        same_lnl(got, ref, 1e-9)
        got5 = pl.evaluate(draws[:5], s.cuda_stream)
        same_lnl(got5, ref[:5], 1e-9)
        # eager work and a synchronisation between the replays
        eager.evaluate(draws[3:6])
        torch.cuda.synchronize()
    eager.close()
    pl.close()


def test_server_results_same_alone_and_in_mix():
    from xpsi_amd import configs
    from xpsi_amd.pipeline import LikelihoodPipeline
    from xpsi_amd.serve import LikelihoodServer
    draws = _draws(64, 40)
    model = configs.config3_model(64)
    alone = LikelihoodPipeline(model, max_batch=1)
    ref = np.array([alone.evaluate([d])[0] for d in draws])
    alone.close()
    say("server: 3 slots x 8, 40 requests in bursts")
    srv = LikelihoodServer(model, n_slots=3, slot_batch=8, use_graphs=True)
    ids = []
    for k, d in enumerate(draws):
        ids.append(srv.submit(d))
        if k % 7 == 6:
            time.sleep(0.002)
    got = {}
    t0 = time.time()
    while len(got) < len(draws) and time.time() - t0 < 60:
        i, l, lat = srv.poll()
        for a, b in zip(i, l):
            got[int(a)] = b
        time.sleep(0.001)
    st = srv.stats()
    srv.close()
    say("server stats %s" % st)
    assert len(got) == len(draws)
    res = np.array([got[i] for i in ids])
    fin = ref > -1e89
    assert np.array_equal(fin, res > -1e89)
    assert np.max(np.abs(res[fin] - ref[fin])) < 1e-9
    assert st["batches"] >= 2


def _config4_requests(n):
    from xpsi_amd import configs
    out = [d for d in configs.config4_draws(4 * n)
           if configs.prior_accepts(d["mass"], d["radius"], 300.0)]
    return out[:n]


def _serve_all(srv, draws, burst=5, pause=0.003, timeout=120.0):
    ids = []
    for k, d in enumerate(draws):
        ids.append(srv.submit(d))
        if k % burst == burst - 1:
            time.sleep(pause)
    got = {}
    t0 = time.time()
    while len(got) < len(draws) and time.time() - t0 < timeout:
        i, l, lat = srv.poll()
        for a, b in zip(i, l):
            got[int(a)] = b
        time.sleep(0.001)
    assert len(got) == len(draws), "server returned %d of %d" % (len(got), len(draws))
    return np.array([got[i] for i in ids])


@pytest.fixture(scope="module")
def config4_alone():
    from xpsi_amd import configs
    from xpsi_amd.pipeline import LikelihoodPipeline
    draws = _config4_requests(20)
    model = configs.config4_model()
    say("config-4 requests alone: %d (sqrt_num_cells %s)"
        % (len(draws), sorted({d["hot_regions"][0]["sqrt_num_cells"] for d in draws})))
    pl = LikelihoodPipeline(model, max_batch=1)
    ref = np.array([pl.evaluate([d])[0] for d in draws])
    pl.close()
    return model, draws, ref


def test_config4_server_alone_vs_mix(config4_alone):
    from xpsi_amd.serve import LikelihoodServer
    model, draws, ref = config4_alone
    say("config-4 server FP64: 3 slots x 8, latency cap 2 ms")
    srv = LikelihoodServer(model, n_slots=3, slot_batch=8, use_graphs=True, max_wait_s=0.002)
    res = _serve_all(srv, draws)
    st = srv.stats()
    srv.close()
    say("config-4 server stats %s" % st)
    fin = ref > -1e89
    assert fin.sum() >= len(draws) // 2
    assert np.array_equal(fin, res > -1e89)
    d = np.max(np.abs(res[fin] - ref[fin]))
    say("config-4 FP64 server vs alone: max |dlnL| %.3e" % d)
    assert d < 1e-9
    assert st["largest_batch"] >= 2 and not st["failed"]
    say("config-4 server precision form")
    srv = LikelihoodServer(dict(model, precision="mixed"), n_slots=3, slot_batch=8,
                           use_graphs=True, max_wait_s=0.002)
    resm = _serve_all(srv, draws)
    srv.close()
    dm = np.max(np.abs(resm[fin] - ref[fin]))
    say("config-4 precision-form server vs FP64 alone: max |dlnL| %.3e" % dm)
    assert dm < 1e-3


def test_config4_admission_budget_binds(config4_alone):
    from xpsi_amd.serve import LikelihoodServer
    model, draws, ref = config4_alone
    srv = LikelihoodServer(model, n_slots=3, slot_batch=8, use_graphs=False)
    fps = np.array([srv.footprint(d) for d in draws])
    srv.close()
    budget = 1.5 * fps.max()
    say("config-4 admission: budget %.3e bytes, footprints %.3e .. %.3e"
        % (budget, fps.min(), fps.max()))
    srv = LikelihoodServer(model, n_slots=3, slot_batch=8, hbm_budget_bytes=budget,
                           use_graphs=False)
    res = _serve_all(srv, draws, burst=len(draws))
    st = srv.stats()
    srv.close()
    say("config-4 admission stats %s" % st)
    fin = ref > -1e89
    assert np.array_equal(fin, res > -1e89)
    assert np.max(np.abs(res[fin] - ref[fin])) < 1e-9
    assert st["peak_inflight_bytes"] <= budget
    assert st["largest_batch"] < 8              # the budget, not the slot, bounds the batches
    assert st["batches"] >= len(draws) // 2


def test_llzero_class_draws_differ_between_calls():
    from xpsi_amd import configs
    from xpsi_amd.pipeline import LikelihoodPipeline
    model = configs.config1()
    bad = dict(mass=1.4, radius=12.0, distance=0.05, cos_inclination=0.5,
               hot_regions=[dict(phase_shift=0.0, super_colatitude=0.8, super_radius=1.0,
                                 super_temperature=6.9)])
    pl = LikelihoodPipeline(model, max_batch=2)
    a = pl.evaluate([bad, bad])
    b = pl.evaluate([bad, bad])
    pl.close()
    say("llzero-class draws: %s then %s" % (a, b))
    for v in np.concatenate([a, b]):
        assert -1e90 <= v <= -1e89
    assert a[0] != b[0] and a[1] != b[1] and a[0] != a[1]
