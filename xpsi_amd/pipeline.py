"""Batched whole-likelihood evaluation on the GPU.

LikelihoodPipeline holds one xh_pipeline (device buffers for `max_batch`
parameter vectors) for a fixed model: instrument response, data, energy grid
and resolution settings.  evaluate() maps a list of parameter vectors to
log-likelihoods; each vector runs mesh construction, ray tables, pulse
integration, energy integration, the response fold and the background
marginalisation on the device (xpsi/Likelihood.py:297-432 for the reference
flow).

Model description (dict):
  spacetime   frequency, star_shape ('AGM_14' | 'sphere')
  hot_regions list of dicts: settings sqrt_num_cells, min/max_sqrt_num_cells,
              num_rays, num_leaves, num_phases, is_antiphased, cede, omit,
              image_order_limit, atm_ext, plus default parameter values
  instrument  matrix [N_ch, N_in], energy_edges [N_in + 1] (keV)
  data        counts [N_ch, N_b], phases [N_b + 1], exposure_time
  num_energies, support [N_ch, 2] or None, marginal settings, llzero,
  interstellar: callable(mid energies) -> attenuation factors, or None
  elsewhere   None or dict: sqrt_num_cells (64), num_rays (1000),
              image_order_limit, elsewhere_temperature (default value)
  precision   'fp64' (default) or 'mixed' (see PRECISIONS)
  fold_terms  mixed-precision response fold: 0 (default) FP32 MFMA with FP64
              slice sums; 6 or 3 bf16 products of a three-term split
  fold_kernel 'auto' (default), 'waves', 'lanes', 'moment' or 'per-point' (see FOLD_KERNELS)
A parameter vector is a dict: mass, radius, distance, cos_inclination and
`hot_regions`: list of dicts phase_shift, super_colatitude, super_radius,
super_temperature, cede_colatitude, cede_radius, cede_azimuth,
cede_temperature, and omit_colatitude, omit_radius, omit_azimuth for a region
configured with omit (HotRegion omit=True: a circular region cut out of the
superseding one); optional per-region sqrt_num_cells; elsewhere_temperature
when the model has an elsewhere component.
"""
import ctypes as C
import math

import numpy as np

from ._lib import lib, check, ModelConfig, RegionConfig, Request, XHError

_REGION_KEYS = ("phase_shift", "super_colatitude", "super_radius", "super_temperature",
                "cede_colatitude", "cede_radius", "cede_azimuth", "cede_temperature")
_OMIT_KEYS = ("omit_colatitude", "omit_radius", "omit_azimuth")


def fill_request(r, p):
    """Copy a parameter-vector dict into an xh_request."""
    r.mass, r.radius = float(p["mass"]), float(p["radius"])
    r.distance, r.cos_inclination = float(p["distance"]), float(p["cos_inclination"])
    for g, hp in enumerate(p["hot_regions"]):
        for q, key in enumerate(_REGION_KEYS):
            v = hp.get(key)
            if v is None:
                v = hp["super_colatitude"] if key == "cede_colatitude" else 0.0
            r.region[g][q] = float(v)
        r.sqrt_num_cells[g] = int(hp.get("sqrt_num_cells", 0))
        r.omit[g][0] = float(hp.get("omit_colatitude", hp["super_colatitude"]))
        r.omit[g][1] = float(hp.get("omit_radius", 0.0))
        r.omit[g][2] = float(hp.get("omit_azimuth", 0.0))
    r.elsewhere_temperature = float(p.get("elsewhere_temperature", 0.0))


def construct_energy_array(num_energies, edges):
    """Energy grid over one instrument waveband (Signal.py:1092-1197)."""
    L, U = edges[0], edges[-1]
    e = np.logspace(np.log10(L), np.log10(U), int(math.ceil(num_energies)), base=10.0)[::-1]
    return np.ascontiguousarray(e[::-1])


# "fp64": FP64 throughout.  "mixed": the precision form - FP32 intensity and
# spline fold on the matrix cores, response fold on FP32 MFMA with FP64 sums
# per 4-deep slice (fold_terms 6 or 3: split-BF16 instead); rays, phase-bin
# accumulation and the likelihood stay FP64.
PRECISIONS = {"fp64": 0, "mixed": 1}
# "moment": spline fold as interval-moment products on the matrix cores;
# "lanes": energies across a wave's lanes, (cell, phase) queries in lockstep;
# "per-point": one spline evaluation per (cell, phase, energy) lane; "auto":
# moment, else lanes, else per-point (autotune_fold measures them).
FOLD_KERNELS = {"auto": 0, "per-point": 1, "moment": 2, "lanes": 3, "waves": 4}


def model_config(model, max_batch=1, max_sqrt_cells=0):
    """Build the C configuration block for a model description; returns
    (config, energies, keep-alive list of the host arrays it points to)."""
    hrs = model["hot_regions"]
    if not 1 <= len(hrs) <= 2:
        raise XHError("one or two hot regions are supported")
    h0 = hrs[0]
    for h in hrs[1:]:
        for key in ("num_rays", "num_leaves", "num_phases", "image_order_limit", "atm_ext"):
            if h.get(key) != h0.get(key):
                raise XHError("hot regions must share %s" % key)
    inst, data = model["instrument"], model["data"]
    energies = construct_energy_array(model["num_energies"], inst["energy_edges"])
    edges = np.ascontiguousarray(inst["energy_edges"], dtype=np.float64)
    matrix = np.ascontiguousarray(inst["matrix"], dtype=np.float64)
    counts = np.ascontiguousarray(data["counts"], dtype=np.float64)
    dph = np.ascontiguousarray(data["phases"], dtype=np.float64)
    support = model.get("support")
    support = None if support is None else np.ascontiguousarray(support, dtype=np.float64)
    att = None
    if model.get("interstellar") is not None:
        mids = 0.5 * (edges[:-1] + edges[1:])
        att = np.ascontiguousarray(model["interstellar"](mids), dtype=np.float64)
    keep = [energies, edges, matrix, counts, dph, support, att]
    cfg = ModelConfig()
    cfg.n_regions = len(hrs)
    for k, h in enumerate(hrs):
        cfg.region[k] = RegionConfig(int(bool(h.get("cede", False))),
                                     int(bool(h.get("is_antiphased", False))),
                                     int(h.get("sqrt_num_cells", 32)),
                                     int(h.get("min_sqrt_num_cells", 10)),
                                     int(h.get("max_sqrt_num_cells", 80)),
                                     int(bool(h.get("omit", False))))
    cfg.num_rays = int(h0.get("num_rays", 200))
    cfg.num_leaves = int(h0.get("num_leaves", 64))
    cfg.num_phases = int(h0.get("num_phases") or cfg.num_leaves)
    cfg.image_order_limit = int(h0.get("image_order_limit") or 0)
    cfg.atm_ext = int(h0.get("atm_ext", 1))
    if cfg.atm_ext == 2:
        atm = model.get("atmosphere")
        if atm is None:
            raise XHError("atm_ext 2 needs model['atmosphere'] = (logT, logg, mu, logE, buf)")
        axes = [np.ascontiguousarray(a, dtype=np.float64) for a in atm[:4]]
        buf = np.ascontiguousarray(atm[4], dtype=np.float64)
        if buf.size != np.prod([len(a) for a in axes]):
            raise XHError("atmosphere buffer size does not match its axes")
        keep += axes + [buf]
        for d in range(4):
            cfg.atm_axes[d] = axes[d].ctypes.data
            cfg.atm_N[d] = len(axes[d])
        cfg.atm_I = buf.ctypes.data
    sp = model["spacetime"]
    cfg.star_shape = {"AGM_14": 0, "sphere": 1}[sp.get("star_shape", "AGM_14")]
    cfg.frequency = float(sp["frequency"])
    cfg.mode_frequency = float(sp.get("mode_frequency", sp["frequency"]))
    cfg.N_E = len(energies)
    cfg.energies = energies.ctypes.data
    cfg.N_in = len(edges) - 1
    cfg.energy_edges = edges.ctypes.data
    cfg.attenuation = None if att is None else att.ctypes.data
    cfg.N_ch = matrix.shape[0]
    cfg.matrix = matrix.ctypes.data
    cfg.N_b = counts.shape[1]
    cfg.data_phases = dph.ctypes.data
    cfg.counts = counts.ctypes.data
    cfg.exposure_time = float(data["exposure_time"])
    cfg.support = None if support is None else support.ctypes.data
    ms = model.get("marginal", {})
    cfg.epsabs = float(ms.get("epsabs", 0.0))
    cfg.epsrel = float(ms.get("epsrel", 1.0e-8))
    cfg.epsilon = float(ms.get("epsilon", 1.0e-3))
    cfg.sigmas = float(ms.get("sigmas", 10.0))
    cfg.slim = float(ms.get("slim", 20.0))
    cfg.llzero = float(model.get("llzero", -1.0e90))
    cfg.max_batch = int(max_batch)
    cfg.max_sqrt_cells = int(max_sqrt_cells)
    prec = model.get("precision", "fp64")
    if prec not in PRECISIONS:
        raise XHError("precision must be one of %s" % sorted(PRECISIONS))
    cfg.precision = PRECISIONS[prec]
    fk = model.get("fold_kernel", "auto")
    if fk not in FOLD_KERNELS:
        raise XHError("fold_kernel must be one of %s" % sorted(FOLD_KERNELS))
    cfg.fold_kernel = FOLD_KERNELS[fk]
    ft = int(model.get("fold_terms", 0))
    if ft not in (0, 3, 6):
        raise XHError("fold_terms must be 0, 3 or 6")
    cfg.fold_terms = ft
    es = model.get("elsewhere")
    if es is not None:
        if es.get("atm_ext", "BB") != "BB":
            raise XHError("the elsewhere component supports the blackbody atmosphere")
        n = int(es.get("sqrt_num_cells", 64))
        if n < 10 or n % 2:
            raise XHError("elsewhere sqrt_num_cells must be an even integer >= 10")
        cfg.elsewhere = 1
        cfg.else_sqrt_num_cells = n
        cfg.else_num_rays = int(es.get("num_rays", 1000))
        cfg.else_image_order_limit = int(es.get("image_order_limit") or 0)
    return cfg, energies, keep


class LikelihoodPipeline:
    def __init__(self, model, max_batch=1, max_sqrt_cells=0):
        cfg, self.energies, self._keep = model_config(model, max_batch, max_sqrt_cells)
        hrs = model["hot_regions"]
        self.model = model
        self.max_batch = int(max_batch)
        self.n_regions = len(hrs)
        self.N_P = cfg.num_phases
        self.N_ch, self.N_b = cfg.N_ch, cfg.N_b
        self._h = lib().xh_pipeline_create(C.byref(cfg))
        if not self._h:
            raise XHError("xh_pipeline_create: " + lib().xh_last_error().decode())
        self._reqs = (Request * self.max_batch)()

    def close(self):
        if getattr(self, "_h", None):
            lib().xh_pipeline_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _fill(self, k, p):
        fill_request(self._reqs[k], p)

    def launch(self, params, stream=None):
        """Start one batch; its results are read by wait().  A pipeline holds
        one batch at a time (its pinned descriptors and results are reused),
        so a second launch before wait() raises."""
        n = len(params)
        if n > self.max_batch:
            raise XHError("batch of %d exceeds max_batch %d" % (n, self.max_batch))
        if getattr(self, "_inflight", False):
            raise XHError("launch: the previous batch has not been waited for")
        for k, p in enumerate(params):
            self._fill(k, p)
        check(lib().xh_pipeline_launch(self._h, n, self._reqs, stream), "xh_pipeline_launch")
        self._n = n
        self._inflight = True

    def wait(self):
        check(lib().xh_pipeline_sync(self._h), "xh_pipeline_sync")
        self._inflight = False
        res = lib().xh_pipeline_results(self._h)
        return np.array([res[k] for k in range(self._n)])

    def evaluate(self, params, stream=None):
        out = []
        for s in range(0, len(params), self.max_batch):
            self.launch(params[s:s + self.max_batch], stream)
            out.append(self.wait())
        return np.concatenate(out)

    def use_graphs(self, on=True):
        check(lib().xh_pipeline_use_graphs(self._h, int(on)), "xh_pipeline_use_graphs")

    def set_timing(self, on=True):
        check(lib().xh_pipeline_set_timing(self._h, int(on)), "xh_pipeline_set_timing")

    STAGES = ("mesh", "rays", "geometry", "profile_fold", "energy_fold", "marginal")

    def stage_ms(self):
        out = np.zeros(6)
        check(lib().xh_pipeline_stage_ms(self._h, out.ctypes.data, 6), "xh_pipeline_stage_ms")
        return dict(zip(self.STAGES, out))

    def work(self, n):
        mpr = lib().xh_pipeline_members_per_request(self._h)
        out = np.zeros(n * mpr)
        check(lib().xh_pipeline_work(self._h, n, out.ctypes.data), "xh_pipeline_work")
        return out.reshape(n, mpr)

    def fetch(self, what, req=0, index=0, size=None):
        names = {"flux": 0, "folded": 1, "sizes": 2, "cellArea": 3, "deflection": 4,
                 "lag": 5, "theta": 6, "phi": 7, "STAR": 8, "else_spectrum": 9}
        size = size or 4 * 1024 * 1024
        buf = np.zeros(size)
        check(lib().xh_pipeline_fetch(self._h, names[what], req, index, buf.ctypes.data, size),
              "xh_pipeline_fetch")
        return buf


def params_from_model(model):
    """The parameter vector stored in a model description."""
    sp = model["spacetime"]
    return dict(mass=sp["mass"], radius=sp["radius"], distance=sp["distance"],
                cos_inclination=sp["cos_inclination"],
                hot_regions=[{k: h.get(k) for k in _REGION_KEYS + _OMIT_KEYS + ("sqrt_num_cells",)
                              if h.get(k) is not None} for h in model["hot_regions"]],
                **({"elsewhere_temperature": model["elsewhere"]["elsewhere_temperature"]}
                   if model.get("elsewhere") is not None else {}))


def autotune_fold(model, params, stream=None, reps=2,
                  candidates=(("wavesx", 1), ("waves", 2), ("waves", 1), ("moment", 4), ("moment", 2),
                              ("moment", 1), ("lanes", 8))):
    """Pick the spline-fold kernel and its block shape for this model and batch.

    Candidates are (kernel, shape): the moment form with blocks of `shape`
    16-phase tiles x `shape` resident 16-energy chunks, or the energy-lane
    form with `shape` phases per wave, or the wave-autonomous moment fold
    with `shape` parallels per wave (precision form, blackbody; "wavesx": one
    block per parallel whose energy chunks share the cell walk).  For each, a pipeline evaluates
    `params` once untimed and `reps` times with HIP-event stage timing; the
    candidate with the least profile-and-fold time wins.  Its shape is set for
    pipelines created afterwards, and the returned model dict selects its
    kernel; every other kernel keeps the best shape measured for it, so a
    later pipeline that cannot use the winner (the wave fold takes the
    blackbody precision form only) still gets a tuned shape.  Returns (model,
    timings) with timings {"kernel/shape": fold ms per evaluation, "chosen":
    "kernel/shape"}."""
    L = lib()

    def select(kernel, shape):
        if kernel in ("waves", "wavesx"):
            check(L.xh_set_wave_fold_rings(int(shape)), "xh_set_wave_fold_rings")
            check(L.xh_set_wave_fold_shared(int(kernel == "wavesx")), "xh_set_wave_fold_shared")
        elif kernel == "moment":
            check(L.xh_set_moment_shape(int(shape)), "xh_set_moment_shape")
        elif kernel == "lanes":
            check(L.xh_set_lane_phases(int(shape)), "xh_set_lane_phases")

    times = {}
    n_e = int(math.ceil(model["num_energies"]))
    for kernel, shape in candidates:
        if kernel == "wavesx" and not 64 < n_e <= 256:
            # the shared walk needs two to four energy chunks: the one-wavefront form instead
            if ("waves", shape) in candidates:
                continue
            kernel = "waves"
        select(kernel, shape)
        try:
            pl = LikelihoodPipeline(dict(model, fold_kernel=kernel.rstrip("x")), max_batch=len(params))
        except XHError:
            continue                      # this kernel does not fit the model
        pl.evaluate(params, stream)
        pl.set_timing(True)
        for _ in range(reps):
            pl.evaluate(params, stream)
        times["%s/%d" % (kernel, shape)] = pl.stage_ms()["profile_fold"] / reps
        pl.close()
    for family in ("waves", "moment", "lanes"):
        own = [k for k in times if k.split("/")[0].rstrip("x") == family]
        if own:
            k_best, s_best = min(own, key=times.get).split("/")
            select(k_best, int(s_best))
    best = min(times, key=times.get)
    kernel, shape = best.split("/")
    select(kernel, int(shape))
    times["chosen"] = best
    return dict(model, fold_kernel=kernel.rstrip("x")), times
