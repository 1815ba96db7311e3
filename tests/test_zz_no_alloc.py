"""GPU: the allocation-free marginal-likelihood entry point against the
workspace one, through the C ABI, on two calls in a row (the second with
fewer channels, so it runs over the scratch the first one left behind)."""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests.models import examples_fast_model

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


def _call(noalloc, T, phases, counts, comps, cph, shifts, pre, sup):
    from xpsi_amd import _device as D
    from xpsi_amd._lib import lib
    L = lib()
    N_ch, N_b = counts.shape
    n_comp, _, N_P = comps.shape
    dc, dph, dcnt = D.to_dev(comps), D.to_dev(phases), D.to_dev(counts)
    dcp, dsh, dpre, dsup = D.to_dev(cph), D.to_dev(shifts), D.to_dev(pre), D.to_dev(sup)
    STAR = D.zeros((N_ch, N_b))
    bml, bmls = D.zeros(N_ch), D.zeros(N_ch)
    lnl = C.c_double(0.0)
    head = (float(T), N_b, D.ptr(dph), N_ch, D.ptr(dcnt), n_comp, D.ptr(dc), N_P, D.ptr(dcp),
            D.ptr(dsh), D.ptr(dpre), D.ptr(dsup), 0.0, 1.0e-8, 1.0e-3, 10.0, -1.0e90, 0, 20.0, None,
            C.byref(lnl), D.ptr(STAR), D.ptr(bml), D.ptr(bmls))
    if noalloc:
        rc = L.xh_eval_marginal_likelihood_noalloc(*head, D.stream_ptr())
    else:
        wsb = L.xh_marginal_workspace_bytes(N_ch)
        ws = D.empty((wsb + 7) // 8)
        rc = L.xh_eval_marginal_likelihood(*head, D.ptr(ws), ws.numel() * 8, D.stream_ptr())
    assert rc == 0, L.xh_last_error().decode()
    D.sync()
    return lnl.value, STAR.cpu().numpy(), bml.cpu().numpy(), bmls.cpu().numpy()


@pytest.fixture(scope="module")
def case(examples_fast):
    model = examples_fast_model(examples_fast)
    _, parts = O.loglikelihood(model, return_parts=True)
    counts = np.ascontiguousarray(model["data"]["counts"], dtype=np.float64)
    comps = np.ascontiguousarray(np.stack(parts["signals"]))
    pre = O.precomputation(counts.astype(np.int32))
    sup = -np.ones((counts.shape[0], 2))
    sup[:, 0] = 0.0
    return counts, comps, pre, sup


def test_noalloc_matches_workspace_entry_point(case):
    counts, comps, pre, sup = case
    phases = np.linspace(0.0, 1.0, 33)
    cph = np.linspace(0.0, 1.0, comps.shape[2])
    shifts = np.array([0.0])
    for n_ch in (counts.shape[0], 97):          # the second call is smaller
        args = (1000.0, phases, counts[:n_ch].copy(), comps[:, :n_ch].copy(), cph, shifts,
                pre[:n_ch].copy(), sup[:n_ch].copy())
        say("marginal, %d channels: workspace entry point" % n_ch)
        ref = _call(False, *args)
        say("marginal, %d channels: allocation-free entry point" % n_ch)
        got = _call(True, *args)
        say("lnL workspace %.12f no-alloc %.12f" % (ref[0], got[0]))
        assert abs(got[0] - ref[0]) < 1.0e-8
        for a, b in zip(got[1:], ref[1:]):
            assert np.max(np.abs(a - b)) <= 1.0e-10 * max(np.max(np.abs(b)), 1e-300)


def test_noalloc_rejects_too_many_channels():
    from xpsi_amd._lib import lib
    L = lib()
    lnl = C.c_double(0.0)
    rc = L.xh_eval_marginal_likelihood_noalloc(1.0, 4, None, 16385, None, 1, None, 8, None, None, None,
                                               None, 0.0, 1e-8, 1e-3, 10.0, -1e90, 0, 20.0, None,
                                               C.byref(lnl), None, None, None, None)
    assert rc == -4
    assert b"16384" in L.xh_last_error()
