"""GPU: the batched atmosphere extension point (xh_atm_init / xh_atm_eval_I /
xh_atm_norm / xh_atm_free, replacing hot_wrapper.pyx:72-252) against the
oracle's sequential restatement of hot_Num4D.pyx:248-438 and the blackbody of
hot_BB.pyx:54-98.  The table queries walk every axis in and out of range, so
the base-node history (clamp on the call that moves to an edge, extrapolate
after) is exercised; the bar is the path's signal bar, 1e-10 relative per query (both
sides run the same FP64 arithmetic; the device's pow/log10 round differently
from the host's by a few ulp, which the cubic weights amplify to ~1e-12)."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests.bars import SIG_RTOL
from tests.test_oracle_pins import BB_NORM, K_B_OVER_KEV, _energy, _poly_table

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


def _queries(ax, n, rng, spill):
    """n queries; a fraction `spill` of them lies up to 0.3 beyond an axis end."""
    q = []
    for a in ax:
        x = rng.uniform(a[0], a[-1], n)
        out = rng.random(n) < spill
        side = rng.random(n) < 0.5
        x[out & side] = a[-1] + rng.uniform(0.0, 0.3, (out & side).sum())
        x[out & ~side] = a[0] - rng.uniform(0.0, 0.3, (out & ~side).sum())
        q.append(x)
    return q


def test_table_sequence_matches_oracle():
    from xpsi_amd.atmosphere import HotAtmosphere
    atm, _ = _poly_table(21, (7, 6, 8, 10))
    rng = np.random.default_rng(22)
    for n, spill in ((1, 0.0), (300, 0.0), (5000, 0.15), (777, 0.6)):
        q = _queries(atm[:4], n, rng, spill)
        # mu beyond [0, 1] is not a query the integrators make
        q[2] = np.clip(q[2], 0.0, 1.0)
        E = _energy(q[0], q[3])
        say("table atmosphere: %d queries, %.0f%% beyond an axis" % (n, 100 * spill))
        ref = O.num4d_sequence(atm, E, q[2], q[0], q[1])
        h = HotAtmosphere(2, atm)
        try:
            got = h.eval_I(E, q[2], np.stack([q[0], q[1]], axis=1))
        finally:
            h.close()
        np.testing.assert_allclose(got, ref, rtol=SIG_RTOL, atol=0.0)


def test_table_norm_and_blackbody():
    from xpsi_amd.atmosphere import HotAtmosphere
    rng = np.random.default_rng(23)
    n = 2000
    lT = rng.uniform(5.5, 7.0, n)
    E = 10.0 ** rng.uniform(-1.5, 1.3, n)
    say("blackbody atmosphere: %d queries" % n)
    h = HotAtmosphere(1)
    try:
        got = h.eval_I(E, rng.uniform(0.0, 1.0, n), np.stack([lT, np.full(n, 14.3)], axis=1))
        assert h.norm() == BB_NORM
    finally:
        h.close()
    kT = K_B_OVER_KEV * 10.0 ** lT
    ref = E ** 3 / np.expm1(E / kT)
    np.testing.assert_allclose(got, ref, rtol=SIG_RTOL)
    atm, _ = _poly_table(24, (5, 5, 5, 5))
    h = HotAtmosphere(2, atm)
    try:
        # hot_Num4D.pyx:441-460: erg / h_keV, erg = 1e-7 J and h_keV the
        # Planck constant in keV s (4.135667662e-18)
        assert h.norm() == pytest.approx(1.0e-7 / 4.135667662e-18, rel=1e-15)
    finally:
        h.close()


def test_bad_extension_raises():
    from xpsi_amd._lib import XHError
    from xpsi_amd.atmosphere import HotAtmosphere
    with pytest.raises(XHError):
        HotAtmosphere(3)
