"""GPU: phase_integrator (xpsi/tools/phase_integrator.pyx:23-121) against the
oracle, element by element to 1e-10 relative, under each phase interpolant
(periodic Akima, Steffen, periodic cubic spline): shifts that wrap and that
do not, a whole-cycle interval, and a negative signal with and without
allow_negative (kept unclipped with it)."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests.bars import SIG_RTOL, elem_relerr

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(params=["Akima", "Steffen", "Cubic"])
def interpolant(request):
    from xpsi_amd.tools.core import set_phase_interpolant
    set_phase_interpolant(request.param)
    O.set_phase_interpolant(request.param)
    yield request.param
    set_phase_interpolant("Akima")
    O.set_phase_interpolant("Akima")


def test_phase_integrator_vs_oracle(interpolant):
    from xpsi_amd.tools.phase_integrator import phase_integrator
    rng = np.random.default_rng(11)
    NP, rows = 48, 7
    sp = np.linspace(0.0, 1.0, NP)
    sig = rng.uniform(0.5, 1.5, (rows, 1)) * (1.2 + np.cos(2.0 * np.pi * (sp[None, :] - rng.uniform(0, 1, (rows, 1)))))
    sig[:, -1] = sig[:, 0]
    for edges in (np.linspace(0.0, 1.0, 33), np.array([0.0, 1.0]), np.sort(rng.uniform(0.0, 1.0, 12))):
        for sh in (0.0, 0.23, -0.61):
            say("%s: %d intervals, shift %.2f" % (interpolant, len(edges) - 1, sh))
            got = phase_integrator(250.0, edges, sig, sp, sh)
            ref = O.phase_integrator(250.0, edges, sig, sp, sh)
            assert elem_relerr(got, ref) < SIG_RTOL
    neg = sig - 1.0   # partly negative rows
    for allow in (False, True):
        got = phase_integrator(250.0, np.linspace(0.0, 1.0, 17), neg, sp, 0.4, allow)
        ref = O.phase_integrator(250.0, np.linspace(0.0, 1.0, 17), neg, sp, 0.4, allow)
        assert elem_relerr(got, ref) < SIG_RTOL
        if allow:
            assert (ref < 0.0).any()
