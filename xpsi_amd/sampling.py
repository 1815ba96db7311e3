"""Sampler-facing callers of the batched likelihood: flat parameter vectors,
many at a time.

The reference's samplers call ``Likelihood.__call__(p)`` once per vector
(xpsi/Likelihood.py:450-513, reached from xpsi/Sample.py:32-130,
xpsi/UltranestSampler.py:79-96 and xpsi/MultinestSampler.py:51-71).
UltraNest (``vectorized=True``) and emcee (``vectorize=True``) can instead
hand the likelihood an array of points ``[n, d]`` and take ``[n]`` values
back.  ``VectorizedLikelihood`` is that callable: the vectors go through one
``LikelihoodPipeline`` launch per ``max_batch`` of them.

Kept from ``Likelihood.__call__``:

- a vector whose log prior is not finite gets ``random_near_llzero`` and no
  device work (Likelihood.py:477-486);
- a log-likelihood at or below ``llzero`` becomes ``random_near_llzero``
  (:505-506), which is ``llzero * (0.1 + 0.9 u)``, u uniform on [0, 1)
  (:268-271); the pipeline already returns such a value where the reference's
  driver does (:305-336), and it is passed on unchanged;
- otherwise the log prior, when a prior is given, is added (:507-511).

``Likelihood`` keeps the reference's one-vector call itself,
``__call__(p=None, reinitialise=False, force=False)``, with its change
detection (a vector equal to the last one returns the stored value unless
``force``), and ``serve_prior_draws`` feeds batches of prior proposals drawn
on the device (``xpsi_amd.prior.DevicePrior``) into the continuous-batching
server.

Vector order: ``parameter_names(model)``.  With one hot region it is the
order of the reference's config-1 style models (mass, radius, distance,
cos_inclination, phase_shift, super_colatitude, super_radius,
super_temperature; the merge order of Spacetime then HotRegion parameters),
with the cede parameters after the super ones when the region has a cede,
and elsewhere_temperature last.  With two regions each region's names carry
its prefix (``p__``, ``s__`` unless the region dict names another).  Pass
``names=`` for any other order.
"""
import numpy as np

from ._lib import XHError
from .pipeline import LikelihoodPipeline

_SUPER = ("phase_shift", "super_colatitude", "super_radius", "super_temperature")
_CEDE = ("cede_colatitude", "cede_radius", "cede_azimuth", "cede_temperature")
_SPACETIME = ("mass", "radius", "distance", "cos_inclination")


def _has_cede(h):
    return bool(h.get("cede")) or h.get("cede_radius") is not None


def _prefixes(model):
    hrs = model["hot_regions"]
    if len(hrs) == 1:
        return [""]
    return [h.get("prefix", "ps"[r]) + "__" for r, h in enumerate(hrs)]


def parameter_names(model):
    """Names of the free parameters, in vector order (see the module docstring)."""
    names = list(_SPACETIME)
    for pre, h in zip(_prefixes(model), model["hot_regions"]):
        names += [pre + k for k in _SUPER + (_CEDE if _has_cede(h) else ())]
    if model.get("elsewhere") is not None:
        names.append("elsewhere_temperature")
    return names


def vector_to_params(model, vec, names=None):
    """The pipeline's parameter dict of one flat vector; resolution settings
    (sqrt_num_cells) come from the model."""
    names = parameter_names(model) if names is None else list(names)
    vec = np.asarray(vec, dtype=np.float64)
    if vec.shape != (len(names),):
        raise XHError("parameter vector has shape %s, expected (%d,)" % (vec.shape, len(names)))
    val = dict(zip(names, vec.tolist()))
    missing = [n for n in parameter_names(model) if n not in val]
    if missing:
        raise XHError("parameter names lack %s" % missing)
    p = {k: val[k] for k in _SPACETIME}
    regions = []
    for pre, h in zip(_prefixes(model), model["hot_regions"]):
        r = {k: val[pre + k] for k in _SUPER}
        if _has_cede(h):
            r.update({k: val[pre + k] for k in _CEDE})
        if h.get("sqrt_num_cells") is not None:
            r["sqrt_num_cells"] = h["sqrt_num_cells"]
        regions.append(r)
    p["hot_regions"] = regions
    if model.get("elsewhere") is not None:
        p["elsewhere_temperature"] = val["elsewhere_temperature"]
    return p


class VectorizedLikelihood:
    """``lnL = f(P)`` for ``P`` of shape ``[n, d]`` (an array back) or ``[d]``
    (a float back), on the device.

    log_prior: optional callable on one vector returning the log prior (the
    reference's ``Prior.__call__``); None means no prior term.
    """

    def __init__(self, model, max_batch=1024, names=None, log_prior=None, seed=None, stream=None):
        self.model = model
        self.names = parameter_names(model) if names is None else list(names)
        vector_to_params(model, np.zeros(len(self.names)), self.names)   # checks the names
        self.max_batch = int(max_batch)
        self.log_prior = log_prior
        self.llzero = float(model.get("llzero", -1.0e90))
        self.stream = stream
        self._rng = np.random.default_rng(seed)
        self._pl = LikelihoodPipeline(model, max_batch=self.max_batch)

    def random_near_llzero(self, n):
        return self.llzero * (0.1 + 0.9 * self._rng.random(n))

    def __call__(self, P):
        P = np.asarray(P, dtype=np.float64)
        single = P.ndim == 1
        P = np.atleast_2d(P)
        if P.ndim != 2 or P.shape[1] != len(self.names):
            raise XHError("points must have shape [n, %d]" % len(self.names))
        n = P.shape[0]
        lp = np.zeros(n)
        if self.log_prior is not None:
            lp = np.array([float(self.log_prior(v)) for v in P])
        ok = np.isfinite(lp)
        out = np.full(n, self.llzero)
        idx = np.flatnonzero(ok)
        if idx.size:
            out[idx] = self._pl.evaluate([vector_to_params(self.model, P[k], self.names) for k in idx],
                                         self.stream)
        # the pipeline reports failed evaluations as random_near_llzero values
        # already (at most 0.1 llzero), which take no prior term either
        redraw = ~ok | (out <= self.llzero)
        out[redraw] = self.random_near_llzero(int(redraw.sum()))
        fin = out > 0.1 * self.llzero
        out[fin] += lp[fin]
        return float(out[0]) if single else out

    def close(self):
        self._pl.close()


class Likelihood:
    """The reference's sampler-facing call, ``Likelihood.__call__(p=None,
    reinitialise=False, force=False) -> float`` (xpsi/Likelihood.py:450-513):
    prior gate, one device evaluation, llzero guard, prior added; a vector
    equal to the previous one returns the stored value unless ``force``
    (the reference's parameter-change detection, ParameterSubspace.py:267-274).
    ``batch(P)`` evaluates many vectors at once (an addition to the reference
    API).  ``p=None`` re-evaluates the current vector."""

    def __init__(self, model, prior=None, names=None, max_batch=1024, seed=None):
        self.model = model
        self.names = parameter_names(model) if names is None else list(names)
        self.prior = prior
        log_prior = None
        if prior is not None:
            log_prior = prior if callable(prior) else None
        self._one = VectorizedLikelihood(model, max_batch=1, names=self.names, log_prior=log_prior, seed=seed)
        self._many = None
        self._max_batch = int(max_batch)
        self._log_prior = log_prior
        self._seed = seed
        self._last_p = None
        self._last = None

    def __call__(self, p=None, reinitialise=False, force=False):
        if reinitialise:
            self._last_p, self._last = None, None
        if p is None:
            if self._last_p is None:
                raise XHError("no parameter vector set")
            p = self._last_p
            force = True
        p = np.asarray(p, dtype=np.float64)
        if not force and self._last_p is not None and np.array_equal(p, self._last_p):
            return self._last
        self._last = float(self._one(p))
        self._last_p = p.copy()
        return self._last

    def batch(self, P):
        if self._many is None:
            self._many = VectorizedLikelihood(self.model, max_batch=self._max_batch, names=self.names,
                                              log_prior=self._log_prior, seed=self._seed)
        return self._many(P)

    def close(self):
        self._one.close()
        if self._many is not None:
            self._many.close()


def serve_prior_draws(server, model, prior, n, rng=None, names=None, batch=512):
    """Draw n proposals from the prior on the device in batches of hypercube
    points, submit those in its support to the server as they come, and
    return (vectors [n, d], request ids [n]).  The results come back through
    ``server.poll()``."""
    names = parameter_names(model) if names is None else list(names)
    rng = np.random.default_rng() if rng is None else rng
    vecs, ids = [], []
    while len(ids) < n:
        P, lp = prior.inverse_sample(rng.random((batch, len(prior))))
        for v in P[np.isfinite(lp)]:
            if len(ids) == n:
                break
            ids.append(server.submit(vector_to_params(model, v, names)))
            vecs.append(v)
    return np.array(vecs), np.array(ids)
