"""GPU: Poisson synthesis of counts (xpsi/tools/synthesise.pyx:40-276).  The
expected counts and scales must match the oracle's phase-bin integrals
exactly (element by element to 1e-10); the realisations are checked as a
distribution: the mean and variance of many bins of known expectation, a
chi-square test of their histogram against the Poisson probabilities, for
expectations on both sides of the inversion/rejection switch at 30, and a
fixed seed reproduces its draws."""
import math
import sys

import numpy as np
import pytest
from scipy import stats

from oracle import oracle as O
from tests.bars import SIG_RTOL, elem_relerr

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu] " + msg, flush=True)
    sys.stdout.flush()


def _case(N_ch=6, N_b=16, NP=48, seed=2):
    rng = np.random.default_rng(seed)
    ph = np.linspace(0.0, 1.0, NP)
    amp = rng.uniform(0.05, 0.5, N_ch)
    comp = amp[:, None] * (1.2 + np.sin(2.0 * math.pi * ph))[None, :]
    bg = rng.uniform(0.5, 2.0, (N_ch, N_b))
    return np.linspace(0.0, 1.0, N_b + 1), comp, ph, bg


def test_expected_counts_and_scales():
    from xpsi_amd.tools.synthesise import synthesise_exposure, synthesise_given_total_count_number
    phases, comp, ph, bg = _case()
    T, nbg = 5000.0, 3.0e4
    E, S, sb = synthesise_exposure(T, phases, [comp], [ph], np.array([0.2]), nbg, bg, gsl_seed=7)
    sb_ref = nbg / bg.sum()
    _, E_ref = O.poisson_likelihood_given_background(T, phases, np.zeros_like(bg), [comp], [ph],
                                                     np.array([0.2]), bg * sb_ref / T)
    say("synthesise_exposure: scale %.12g (oracle %.12g), EXPEC rel err %.3e" % (sb, sb_ref, elem_relerr(E, E_ref)))
    assert sb == pytest.approx(sb_ref, rel=1e-14)
    assert elem_relerr(E, E_ref) < SIG_RTOL
    assert np.all(S == np.round(S)) and np.all(S >= 0)
    E2, S2, ss, ratio = synthesise_given_total_count_number(phases, 2.0e5, [comp], [ph], np.array([0.2]), nbg,
                                                            bg, gsl_seed=7)
    _, star = O.poisson_likelihood_given_background(1e5, phases, np.zeros_like(bg), [comp], [ph],
                                                    np.array([0.2]), np.zeros_like(bg))
    ss_ref = 2.0e5 / star.sum()
    E2_ref = star * ss_ref + bg * sb_ref
    assert ss == pytest.approx(ss_ref, rel=1e-12)
    assert ratio == pytest.approx(sb_ref / ss_ref, rel=1e-12)
    assert elem_relerr(E2, E2_ref) < SIG_RTOL
    assert float(E2.sum()) == pytest.approx(2.0e5 + nbg, rel=1e-12)
    # a fixed seed reproduces its realisation; another seed does not
    _, S_again, _ = synthesise_exposure(T, phases, [comp], [ph], np.array([0.2]), nbg, bg, gsl_seed=7)
    _, S_other, _ = synthesise_exposure(T, phases, [comp], [ph], np.array([0.2]), nbg, bg, gsl_seed=8)
    assert np.array_equal(S, S_again) and not np.array_equal(S, S_other)


@pytest.mark.parametrize("lam", [0.7, 5.0, 29.5, 30.5, 140.0, 2.5e4])
def test_poisson_draws_distribution(lam):
    from xpsi_amd.tools.synthesise import synthesise_exposure
    # a flat, zero star signal and a uniform background: every bin expects lam
    N_ch, N_b = 200, 200
    phases = np.linspace(0.0, 1.0, N_b + 1)
    ph = np.linspace(0.0, 1.0, 8)
    zero = np.zeros((N_ch, 8))
    bg = np.ones((N_ch, N_b))
    E, S, _ = synthesise_exposure(1.0, phases, [zero], [ph], np.array([0.0]), lam * N_ch * N_b, bg,
                                  gsl_seed=1234 + int(lam * 10))
    np.testing.assert_allclose(E, lam, rtol=1e-12)
    x = S.reshape(-1)
    n = x.size
    m, v = x.mean(), x.var(ddof=1)
    say("lambda %g: mean %.5g var %.5g over %d draws" % (lam, m, v, n))
    assert abs(m - lam) < 5.0 * math.sqrt(lam / n)
    assert abs(v - lam) < 6.0 * lam * math.sqrt(2.0 / n) + 6.0 * math.sqrt(lam / n)
    # chi-square of the histogram against the Poisson probabilities (cells
    # merged until each expects at least 20 draws)
    lo, hi = int(stats.poisson.ppf(1e-6, lam)), int(stats.poisson.ppf(1 - 1e-6, lam))
    edges, obs, exp = [], [], []
    acc_o = acc_e = 0.0
    for k in range(lo, hi + 1):
        acc_o += np.count_nonzero(x == k)
        acc_e += n * stats.poisson.pmf(k, lam)
        if acc_e >= 20.0:
            obs.append(acc_o); exp.append(acc_e); acc_o = acc_e = 0.0
    obs[-1] += acc_o + np.count_nonzero(x > hi) + np.count_nonzero(x < lo)
    exp[-1] += acc_e + n * (stats.poisson.sf(hi, lam) + stats.poisson.cdf(lo - 1, lam))
    chi2 = float(np.sum((np.array(obs) - np.array(exp)) ** 2 / np.array(exp)))
    p = stats.chi2.sf(chi2, len(obs) - 1)
    say("lambda %g: chi2 %.1f over %d cells, p %.3g" % (lam, chi2, len(obs), p))
    assert p > 1e-4
