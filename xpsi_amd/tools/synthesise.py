"""Drop-in for xpsi/tools/synthesise.pyx: synthesise_exposure (:40-148) and
synthesise_given_total_count_number (:152-276), with the reference
signatures and return tuples.  The expected counts follow the reference
exactly; the Poisson realisation is drawn on the device from ``gsl_seed``
(a clock-based seed when None) with the same distribution but not GSL's
generator, so a given seed does not reproduce the reference's numbers.
Components must share one phase grid."""
import ctypes as C
import time

import numpy as np

from .._lib import lib, check
from .. import _device as D


def _prepare(phases, components, component_phases, phase_shifts, background, allow_negative):
    ph0 = np.asarray(component_phases[0], dtype=np.float64)
    for cp in component_phases[1:]:
        if not np.array_equal(np.asarray(cp), ph0):
            raise NotImplementedError("components must share one phase grid")
    if not isinstance(allow_negative, bool):
        if len(set(bool(a) for a in allow_negative)) != 1:
            raise NotImplementedError("per-component allow_negative is not supported")
        allow_negative = bool(allow_negative[0])
    comps = np.ascontiguousarray(np.stack([np.asarray(c, dtype=np.float64) for c in components]))
    n_comp, N_ch, N_P = comps.shape
    N_b = len(phases) - 1
    bg = np.ascontiguousarray(background, dtype=np.float64)
    if bg.shape != (N_ch, N_b):
        raise TypeError("the background must have one entry per channel and phase interval")
    dev = dict(c=D.to_dev(comps), ph=D.to_dev(phases), cp=D.to_dev(ph0),
               sh=D.to_dev(np.asarray(phase_shifts, dtype=np.float64)), bg=D.to_dev(bg))
    return dev, n_comp, N_ch, N_P, N_b, int(allow_negative)


def _seed(gsl_seed):
    return int(time.time_ns() if gsl_seed is None else gsl_seed) & 0xFFFFFFFFFFFFFFFF


def synthesise_exposure(exposure_time, phases, components, component_phases, phase_shifts,
                        expected_background_counts, background, allow_negative=False, gsl_seed=None):
    dev, n_comp, N_ch, N_P, N_b, an = _prepare(phases, components, component_phases, phase_shifts,
                                               background, allow_negative)
    L = lib()
    ws = D.empty((L.xh_synthesise_workspace_bytes(N_ch, N_b) + 7) // 8)
    EXPEC, SYN = D.empty((N_ch, N_b)), D.empty((N_ch, N_b))
    sb = C.c_double(0.0)
    check(L.xh_synthesise_exposure(float(exposure_time), N_b, D.ptr(dev["ph"]), N_ch, n_comp, D.ptr(dev["c"]),
                                   N_P, D.ptr(dev["cp"]), D.ptr(dev["sh"]), float(expected_background_counts),
                                   D.ptr(dev["bg"]), an, _seed(gsl_seed), D.ptr(EXPEC), D.ptr(SYN), C.byref(sb),
                                   D.ptr(ws), ws.numel() * 8, D.stream_ptr()), "xh_synthesise_exposure")
    return EXPEC.cpu().numpy(), SYN.cpu().numpy(), sb.value


def synthesise_given_total_count_number(phases, expected_star_counts, components, component_phases,
                                        phase_shifts, expected_background_counts, background,
                                        allow_negative=False, gsl_seed=None):
    dev, n_comp, N_ch, N_P, N_b, an = _prepare(phases, components, component_phases, phase_shifts,
                                               background, allow_negative)
    L = lib()
    ws = D.empty((L.xh_synthesise_workspace_bytes(N_ch, N_b) + 7) // 8)
    EXPEC, SYN = D.empty((N_ch, N_b)), D.empty((N_ch, N_b))
    sc = np.zeros(2)
    check(L.xh_synthesise_given_total_count_number(N_b, D.ptr(dev["ph"]), float(expected_star_counts), N_ch,
                                                   n_comp, D.ptr(dev["c"]), N_P, D.ptr(dev["cp"]),
                                                   D.ptr(dev["sh"]), float(expected_background_counts),
                                                   D.ptr(dev["bg"]), an, _seed(gsl_seed), D.ptr(EXPEC),
                                                   D.ptr(SYN), sc.ctypes.data, D.ptr(ws), ws.numel() * 8,
                                                   D.stream_ptr()),
          "xh_synthesise_given_total_count_number")
    return EXPEC.cpu().numpy(), SYN.cpu().numpy(), float(sc[0]), float(sc[1])
