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
        assert np.array_equal(got, ref) or np.max(np.abs(got - ref)) < 1e-9
        got5 = pl.evaluate(draws[:5], s.cuda_stream)
        assert np.max(np.abs(got5 - ref[:5])) < 1e-9
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
