"""GPU batched form: prior draws of the batched workload (polar caps, compact
stars, uneven mesh and energy resolutions) against the oracle, and each
result alone versus in the mix."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests.batch_cases import sample, oracle_model

pytestmark = pytest.mark.gpu
LNL_ATOL = 1.0e-8
LLZERO = -1.0e90


def say(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


def classify(x):
    return "zero" if x <= 0.1 * LLZERO else "finite"


@pytest.fixture(scope="module")
def cases():
    cs = sample(16)
    ref = [O.loglikelihood(oracle_model(d, d["num_energies"])) for d in cs]
    return cs, ref


def test_batched_vs_oracle_and_alone(cases):
    from xpsi_amd import configs
    from xpsi_amd.pipeline import LikelihoodPipeline
    cs, ref = cases
    groups = {}
    for k, d in enumerate(cs):
        groups.setdefault(d["num_energies"], []).append(k)
    for ne, ks in sorted(groups.items()):
        say("batch of %d draws, %d energies" % (len(ks), ne))
        pl = LikelihoodPipeline(configs.config3_model(ne), max_batch=len(ks))
        got = pl.evaluate([cs[k] for k in ks])
        pl1 = LikelihoodPipeline(configs.config3_model(ne), max_batch=1)
        for g, k in zip(got, ks):
            alone = pl1.evaluate([cs[k]])[0]
            say("  draw %d: batch %.10g alone %.10g oracle %.10g" % (k, g, alone, ref[k]))
            assert classify(g) == classify(ref[k])
            assert classify(alone) == classify(g)
            if classify(ref[k]) == "finite":
                assert abs(g - ref[k]) < LNL_ATOL
                assert abs(alone - g) < 1.0e-9
        pl.close()
        pl1.close()


def test_groups_on_concurrent_streams(cases):
    """Every energy group's pipeline on its own non-blocking stream, all
    launched before any is waited on (how the batched workload overlaps its
    groups): the same lnL as one group at a time, and the same fold work count
    (read back on each pipeline's own stream)."""
    import torch
    from xpsi_amd import configs
    from xpsi_amd.pipeline import LikelihoodPipeline
    cs, ref = cases
    groups = {}
    for k, d in enumerate(cs):
        groups.setdefault(d["num_energies"], []).append(k)
    pls = {ne: LikelihoodPipeline(configs.config3_model(ne), max_batch=len(ks)) for ne, ks in groups.items()}
    serial = {ne: pls[ne].evaluate([cs[k] for k in ks]) for ne, ks in groups.items()}
    work_serial = {ne: pls[ne].work(len(ks)) for ne, ks in groups.items()}
    streams = {ne: torch.cuda.Stream() for ne in groups}
    for rep in range(2):
        for ne in sorted(groups, reverse=True):
            pls[ne].launch([cs[k] for k in groups[ne]], streams[ne].cuda_stream)
        got = {ne: pls[ne].wait() for ne in groups}
        for ne, ks in sorted(groups.items()):
            w = pls[ne].work(len(ks))
            say("concurrent pass %d, %d energies: max |lnL - serial| %.3e, work %s" %
                (rep, ne, float(np.max(np.abs(got[ne] - serial[ne]))), np.array_equal(w, work_serial[ne])))
            assert np.array_equal(w, work_serial[ne])
            for g, s, k in zip(got[ne], serial[ne], ks):
                assert classify(g) == classify(s)
                if classify(s) == "finite":
                    assert abs(g - s) < 1.0e-9
                    assert abs(g - ref[k]) < LNL_ATOL
    torch.cuda.synchronize()
    for pl in pls.values():
        pl.close()
