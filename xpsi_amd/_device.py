"""Device buffers for the drop-in entry points (PyTorch-ROCm as allocator)."""
import numpy as np


def torch():
    import torch as _t
    return _t


def require_gpu():
    t = torch()
    if not t.cuda.is_available():
        from ._lib import XHError
        raise XHError("no ROCm device visible: the xpsi_amd kernels need an MI355X")
    return t


def to_dev(a, dtype=np.float64):
    t = require_gpu()
    arr = np.ascontiguousarray(a, dtype=dtype)
    return t.from_numpy(arr).to("cuda")


def empty(shape, dtype="float64"):
    t = require_gpu()
    return t.empty(shape, dtype=getattr(t, dtype), device="cuda")


def zeros(shape, dtype="float64"):
    t = require_gpu()
    return t.zeros(shape, dtype=getattr(t, dtype), device="cuda")


def ptr(x):
    return None if x is None else x.data_ptr()


def stream_ptr():
    t = torch()
    return t.cuda.current_stream().cuda_stream


def sync():
    torch().cuda.synchronize()
