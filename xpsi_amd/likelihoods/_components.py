"""Component signals for the device likelihoods, as the reference takes
them (xpsi/tools/compute_expected_counts.pyx:26-63, :66-197): each component
a [channels, phases] array on its own phase grid, and allow_negative either
one bool for every component or one per component."""
import numpy as np

from .. import _device as D


def allow_flags(allow_negative, n_comp):
    """build_allow_negative_array: a bool applies to every component, else one
    flag per component (ValueError on a count mismatch)."""
    if isinstance(allow_negative, (bool, np.bool_)):
        return [bool(allow_negative)] * n_comp
    flags = [bool(a) for a in allow_negative]
    if len(flags) != n_comp:
        raise ValueError('Number of allow_negative declarations does not match the number of components.')
    return flags


def pack(components, component_phases, N_ch):
    """Returns (shared, n_comp, N_P, comps_dev, phases_dev, nphase_dev): with
    one shared grid, the [comp][channel][phase] block and that grid (nphase_dev
    None); otherwise the components' [channel][phase] blocks one after another,
    their grids likewise and the phases per component."""
    comps = [np.ascontiguousarray(c, dtype=np.float64) for c in components]
    grids = [np.ascontiguousarray(p, dtype=np.float64) for p in component_phases]
    if len(comps) != len(grids):
        raise TypeError("one phase grid per component")
    for c, g in zip(comps, grids):
        if c.ndim != 2 or c.shape[0] != N_ch or c.shape[1] != len(g):
            raise TypeError("each component must be [channels, its phases]")
    shared = all(len(g) == len(grids[0]) and np.array_equal(g, grids[0]) for g in grids)
    if shared:
        return (True, len(comps), len(grids[0]), D.to_dev(np.stack(comps)), D.to_dev(grids[0]), None)
    flat = np.concatenate([c.ravel() for c in comps])
    return (False, len(comps), 0, D.to_dev(flat), D.to_dev(np.concatenate(grids)),
            D.to_dev(np.array([len(g) for g in grids], dtype=np.int32), np.int32))
