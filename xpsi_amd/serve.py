"""Continuous-batching likelihood server (wrapper of xh_server_*).

A nested sampler or MCMC driver submits parameter vectors as they are
proposed; the native dispatcher admits them into free pipeline slots (each
with its own stream and device buffers) while the in-flight device footprint
stays within the budget, so batches overlap on the GPU.  Results come back
with their submit-to-result latency.  The reference evaluates one vector per
call (xpsi/Likelihood.py:450-513); this interface is additive.
"""
import ctypes as C

import numpy as np

from ._lib import lib, Request, XHError
from .pipeline import model_config, fill_request


class LikelihoodServer:
    """n_slots pipelines of slot_batch requests each.  hbm_budget_bytes bounds
    the device footprint of the batches in flight (<= 0: unbounded);
    max_wait_s is how long a partial batch may wait for more requests while
    another batch runs (0: dispatch at once)."""

    def __init__(self, model, n_slots=4, slot_batch=64, hbm_budget_bytes=0.0,
                 use_graphs=True, max_sqrt_cells=0, max_wait_s=0.0):
        cfg, self.energies, self._keep = model_config(model, slot_batch, max_sqrt_cells)
        self._h = lib().xh_server_create(C.byref(cfg), int(n_slots), int(slot_batch),
                                         float(hbm_budget_bytes), int(use_graphs))
        if not self._h:
            raise XHError("xh_server_create: " + lib().xh_last_error().decode())
        if lib().xh_server_set_policy(self._h, float(max_wait_s)) != 0:
            raise XHError("xh_server_set_policy: bad max_wait_s")
        self._req = Request()
        self._id = C.c_longlong(0)

    def footprint(self, params):
        """Device bytes the admission control charges for one request."""
        fill_request(self._req, params)
        return lib().xh_server_footprint(self._h, C.byref(self._req))

    def warm(self):
        """Capture every slot's graph for every batch size up front."""
        if lib().xh_server_warm(self._h) != 0:
            raise XHError("xh_server_warm: " + lib().xh_last_error().decode())

    def submit(self, params):
        fill_request(self._req, params)
        lib().xh_server_submit(self._h, C.byref(self._req), C.byref(self._id))
        return self._id.value

    def poll(self, max_n=4096):
        ids = np.zeros(max_n, dtype=np.int64)
        lnl = np.zeros(max_n)
        lat = np.zeros(max_n)
        n = lib().xh_server_poll(self._h, ids.ctypes.data, lnl.ctypes.data, lat.ctypes.data, max_n)
        return ids[:n], lnl[:n], lat[:n]

    def stats(self):
        out = np.zeros(8)
        lib().xh_server_stats(self._h, out.ctypes.data, 8)
        return dict(queued=out[0], inflight_bytes=out[1], batches=out[2], completed=out[3],
                    mean_fill=out[4], peak_inflight_bytes=out[5], largest_batch=int(out[6]),
                    failed=bool(out[7]))

    def close(self):
        if getattr(self, "_h", None):
            lib().xh_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
