"""The prior on the device: inverse sampling of many hypercube points at once.

Follows xpsi/Prior.py:84-160 (flat densities between each parameter's
bounds: ``inverse_sample`` maps a coordinate u to lo + (hi - lo) u and
``__call__`` is 0 inside the bounds, -inf outside) together with the
compactness cuts of the reference's example prior
(examples/examples_fast/Modules/CustomPrior.py:49-66: radius <= 16 km, the
polar radius above the photon-sphere bound, one maximum of |z| on the
surface), which use the spacetime quantities of xpsi/Spacetime.py:111-190.
A nested sampler's or MCMC driver's batch of unit-hypercube points becomes
parameter vectors and log priors in one launch; the accepted vectors go on
to the likelihood (``VectorizedLikelihood``, ``LikelihoodServer``).
"""
import numpy as np

from ._lib import lib, check, XHError
from . import _device as D


class DevicePrior:
    """bounds: sequence of (lo, hi) per vector coordinate, in vector order;
    names: the coordinates' names (``mass`` and ``radius``, when present,
    switch the compactness cuts on); frequency: the star's spin (Hz)."""

    def __init__(self, bounds, names, frequency, compactness=True):
        self.bounds = np.ascontiguousarray(bounds, dtype=np.float64)
        if self.bounds.ndim != 2 or self.bounds.shape[1] != 2:
            raise XHError("bounds must have shape [d, 2]")
        self.names = list(names)
        if len(self.names) != self.bounds.shape[0]:
            raise XHError("one name per bound")
        if not np.all(np.isfinite(self.bounds)):
            raise ValueError("Compact support required.")
        self.frequency = float(frequency)
        self.idx_mass = self.names.index("mass") if (compactness and "mass" in self.names) else -1
        self.idx_radius = self.names.index("radius") if (compactness and "radius" in self.names) else -1
        self._dbounds = None

    def __len__(self):
        return self.bounds.shape[0]

    def inverse_sample(self, hypercube):
        """hypercube [n, d] (numpy or a device tensor) -> (params [n, d],
        log prior [n]) as numpy arrays."""
        t = D.require_gpu()
        if self._dbounds is None:
            self._dbounds = D.to_dev(self.bounds)
        if isinstance(hypercube, t.Tensor):
            hc = hypercube.to(device="cuda", dtype=t.float64).contiguous()
        else:
            hc = D.to_dev(np.atleast_2d(hypercube))
        n, d = hc.shape
        if d != len(self):
            raise XHError("hypercube points must have %d coordinates" % len(self))
        params, lp = D.empty((n, d)), D.empty(n)
        check(lib().xh_prior_inverse_sample(n, d, D.ptr(hc), D.ptr(self._dbounds), self.idx_mass,
                                            self.idx_radius, self.frequency, D.ptr(params), D.ptr(lp),
                                            D.stream_ptr()), "xh_prior_inverse_sample")
        return params.cpu().numpy(), lp.cpu().numpy()

    def draw(self, n, rng=None):
        """n draws from the prior (rejecting those outside its support), with
        the hypercube points from rng (numpy Generator)."""
        rng = np.random.default_rng() if rng is None else rng
        out = []
        while sum(len(o) for o in out) < n:
            p, lp = self.inverse_sample(rng.random((max(2 * n, 64), len(self))))
            out.append(p[np.isfinite(lp)])
        return np.concatenate(out)[:n]
