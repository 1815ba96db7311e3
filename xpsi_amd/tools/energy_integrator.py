"""Drop-in for xpsi/tools/energy_integrator.pyx:27-114."""
import numpy as np

from .._lib import lib, check
from .. import _device as D


def energy_integrator(N_Ts, signal, energies, energy_edges):
    """Integrate signal[N_E, N_P] over log10-energy intervals.

    `energies` and `energy_edges` are log10 values, as in the reference.
    Returns [N_edges - 1, N_P].
    """
    sig = D.to_dev(signal)
    N_E, N_P = sig.shape
    le, ed = D.to_dev(energies), D.to_dev(energy_edges)
    out = D.empty((len(energy_edges) - 1, N_P))
    rc = lib().xh_energy_integrator(N_E, N_P, D.ptr(sig), D.ptr(le), len(energy_edges),
                                    D.ptr(ed), D.ptr(out), D.stream_ptr())
    check(rc, "xh_energy_integrator")
    return out.cpu().numpy()
