"""Agreement measures shared by the GPU parity tests.

The path's bars: signal arrays to 1e-10 relative, lnL to 1e-8 absolute (FP64).
Signal arrays are compared element by element: each element's error is taken
relative to its own magnitude, or to FLOOR times the array's largest magnitude
where the element is smaller than that (elements that small are zero or sit
at the edge of visibility, where both implementations round a near-cancelled
value).
"""
import numpy as np

SIG_RTOL = 1.0e-10
LNL_ATOL = 1.0e-8
FLOOR = 1.0e-6


def elem_relerr(a, b, floor=FLOOR):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    big = float(np.max(np.abs(b))) if b.size else 0.0
    if big == 0.0:
        return float(np.max(np.abs(a))) if a.size else 0.0
    scale = np.maximum(np.abs(b), floor * big)
    return float(np.max(np.abs(a - b) / scale))


def norm_relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
