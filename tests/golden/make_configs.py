"""Writes synthetic data fixtures for the benchmark workloads with the oracle.

python tests/golden/make_configs.py [config1] [config2] [config5] [config5_full]

config1.npz: counts[100, 32] ~ Poisson(exposure * phase-binned star rate + 5)
with numpy.random.default_rng(42); the rate is the oracle's folded signal of
xpsi_amd.configs.config1 integrated over each phase bin (the expected-count
integral of compute_expected_counts.pyx:66-197).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from xpsi_amd import configs  # noqa: E402


def binned_rate(signal, comp_phases, edges, shift):
    out = np.zeros((signal.shape[0], len(edges) - 1))
    for c in range(signal.shape[0]):
        sp = O.akima(comp_phases, signal[c], periodic=True)
        for j in range(len(edges) - 1):
            pa, pb = edges[j] + shift, edges[j + 1] + shift
            pa -= np.floor(pa)
            pb -= np.floor(pb)
            if abs((edges[j + 1] - edges[j]) - 1.0) < 1e-12:
                pa, pb = 0.0, 1.0
            if pa < pb:
                v = O.akima_integ(sp, pa, pb)
            else:
                v = O.akima_integ(sp, pa, 1.0) + O.akima_integ(sp, 0.0, pb)
            out[c, j] = max(v, 0.0)
    return out


def make_config1():
    model = configs.config1(counts=np.zeros((100, 32)))
    lnl, parts = O.loglikelihood(model, return_parts=True)
    sig = parts["signals"][0]
    rate = binned_rate(sig, np.linspace(0.0, 1.0, 64), model["data"]["phases"], 0.0)
    expected = model["data"]["exposure_time"] * rate + 5.0
    rng = np.random.default_rng(42)
    counts = rng.poisson(expected).astype(np.float64)
    np.savez_compressed(os.path.join(HERE, "config1.npz"), counts=counts, expected=expected)
    print("config1: total counts %.4g" % counts.sum())


def make_config2(atm):
    """config2_<atm>.npz: counts[300, 32] from both hot regions, default_rng(42)."""
    model = configs.config2(atm, counts=np.zeros((300, 32)))
    lnl, parts = O.loglikelihood(model, return_parts=True)
    rate = np.zeros((300, 32))
    for sig, h in zip(parts["signals"], model["hot_regions"]):
        rate += binned_rate(sig, np.linspace(0.0, 1.0, 128), model["data"]["phases"], h["phase_shift"])
    expected = model["data"]["exposure_time"] * rate + 5.0
    rng = np.random.default_rng(42)
    counts = rng.poisson(expected).astype(np.float64)
    np.savez_compressed(os.path.join(HERE, "config2_%s.npz" % atm), counts=counts, expected=expected)
    print("config2 %s: total counts %.4g" % (atm, counts.sum()))


def make_config5(scale):
    """config5_<scale>.npz: counts from both hot regions (with the elsewhere
    spectrum and interstellar absorption), default_rng(42)."""
    model = configs.config5(scale)
    n_ch, n_b = model["data"]["counts"].shape
    lnl, parts = O.loglikelihood(model, return_parts=True)
    rate = np.zeros((n_ch, n_b))
    n_p = model["hot_regions"][0]["num_leaves"]
    for sig, h in zip(parts["signals"], model["hot_regions"]):
        rate += binned_rate(sig, np.linspace(0.0, 1.0, n_p), model["data"]["phases"], h["phase_shift"])
    expected = model["data"]["exposure_time"] * rate + 5.0
    rng = np.random.default_rng(42)
    counts = rng.poisson(expected).astype(np.float64)
    np.savez_compressed(os.path.join(HERE, "config5_%s.npz" % scale), counts=counts, expected=expected)
    print("config5 %s: total counts %.4g" % (scale, counts.sum()))


if __name__ == "__main__":
    which = sys.argv[1:] or ["config1", "config2", "config5"]
    if "config1" in which:
        make_config1()
    if "config2" in which:
        make_config2("BB")
        make_config2("Num4D")
    if "config5" in which:
        make_config5("test")
    if "config5_full" in which:
        make_config5("full")
