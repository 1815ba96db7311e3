"""The C-ABI library loads on the CPU and exports every entry point that
include/xpsi_hip.h declares (no compute call without a GPU)."""
import os
import re

from xpsi_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "xpsi_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(xh_[a-z0-9_]+)\s*\(", src)))


def test_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(_lib.EXPORTED) >= set(names) - set(), sorted(set(names) - set(_lib.EXPORTED))


def test_no_gpu_means_loud_failure():
    import numpy as np
    import torch
    if torch.cuda.is_available():
        return
    from xpsi_amd.tools import energy_integrator
    try:
        energy_integrator(1, np.ones((8, 2)), np.linspace(0, 1, 8), np.linspace(0, 1, 4))
    except _lib.XHError:
        pass
    else:
        raise AssertionError("drop-in ran without a GPU")
