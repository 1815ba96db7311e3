"""Drop-in for xpsi/tools/phase_integrator.pyx.

phase_integrator (:23-121) keeps the reference signature and returns the
[rows, intervals] array: each row of the signal, interpolated over its phases
with the process-wide phase interpolant (xpsi_amd.tools.core), integrated
over the phase intervals shifted by phase_shift, each piece kept where
positive or allow_negative, times the exposure time.
"""
import numpy as np

from .._lib import lib, check
from .. import _device as D


def phase_integrator(exposure_time, phases, signal, signal_phases, phase_shift, allow_negative=0):
    signal = np.ascontiguousarray(signal, dtype=np.float64)
    if signal.ndim != 2:
        raise TypeError("the signal must be [rows, phases]")
    N_rows, N_P = signal.shape
    phases = np.ascontiguousarray(phases, dtype=np.float64)
    signal_phases = np.ascontiguousarray(signal_phases, dtype=np.float64)
    if len(signal_phases) != N_P:
        raise TypeError("one signal phase per signal column")
    N_b = len(phases) - 1
    dsig, dph, dsp = D.to_dev(signal), D.to_dev(phases), D.to_dev(signal_phases)
    out = D.empty((N_rows, N_b))
    check(lib().xh_phase_integrator(float(exposure_time), N_b, D.ptr(dph), N_rows, D.ptr(dsig), N_P, D.ptr(dsp),
                                    float(phase_shift), int(bool(allow_negative)), D.ptr(out), D.stream_ptr()),
          "xh_phase_integrator")
    return out.cpu().numpy()
