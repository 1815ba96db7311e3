"""Benchmark: batched lnL evaluations/s on one MI355X.

Workload (the batched form): 4096 parameter vectors drawn from the prior with
uneven resolutions (sqrt_num_cells in {24,32,48,64,96}, num_energies in
{64,128,256}), one blackbody hot spot, 100-channel Gaussian response, 32
phase bins; vectors failing the compactness prior are answered without
evaluation and not counted.  A step evaluates every accepted vector once
(one launch per energy count).  The timed form is the precision form by
default (--precision mixed); its |dlnL| against an FP64 pass over the same
vectors is reported beside the rate.  Also reported: the plain FP64 GPU path
one request at a time (config-1 request) with p50/p99 latency, the same
config-1 request batched, Poisson-arrival serving, and the CPU oracle on the
host cores.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision mixed|fp64] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
FP32_MFMA_PEAK_TFLOPS = 157.3
KERNEL_NAMES = {"moment": "k_moment_fold", "lanes": "k_lane_fold", "waves": "k_wave_fold",
                "wavesx": "k_wave_fold_shared"}
# Bytes per fold launch on the default batched workload (4096 prior draws,
# precision form) from rocprofv3 --pmc FETCH_SIZE (x 2.0) and WRITE_SIZE (x 1.0;
# factors from a 512 MiB device copy), KiB per launch for the 256-, 128- and
# 64-energy launches of one step, per autotuned fold (profiles/wave_fold_batched_default.md,
# profiles/moment_fold_batched_default.md hold the runs).
# The wave fold's writes are its FP64 atomic additions into the flux (one set
# per accumulation chain of 8 intervals); its reads include the reloads of its
# register spills (240 bytes per lane of scratch at four waves per SIMD).
FOLD_TRAFFIC_KIB = {
    "waves/1": {256: (1692562, 31260897), 128: (834653, 13825472), 64: (527687, 7511708)},
    "wavesx/1": {256: (122594, 27052533), 128: (111036, 11906776)},
    "moment/2": {256: (17123810, 20011418), 128: (9762315, 9300707), 64: (6796957, 5185836)},
}
HBM_PEAK_GBS = 8000.0


def log(msg):
    """Progress on stderr (the JSON line on stdout stays the only output there)."""
    print("[bench %7.1fs] %s" % (time.perf_counter() - _T0, msg), file=sys.stderr, flush=True)


_T0 = time.perf_counter()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--single", type=int, default=40, help="plain single-request evaluations")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--fold-ec", type=int, default=0, help="energies per fold block (0 = auto)")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--one-stream", action="store_true", help="launch every energy group on one stream")
    ap.add_argument("--no-graphs", action="store_true", help="launch the batched step's kernels eagerly")
    ap.add_argument("--fold", default="auto",
                    help="spline fold of the batched form: auto (autotuned per energy group) or kernel/shape, "
                         "e.g. waves/2 or moment/2")
    ap.add_argument("--wave-chain", type=int, default=-1,
                    help="intervals per FP32 chain of the wave fold (-1 = library default, 0 = per parallel)")
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-batch", type=int, default=256)
    ap.add_argument("--serve-wait", type=float, default=0.002, help="latency cap of a partial batch (s)")
    ap.add_argument("--serve-budget", type=float, default=200e9, help="device footprint budget (bytes)")
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--config1-batch", type=int, default=1024)
    ap.add_argument("--c5-n", type=int, default=1024, help="large-workload vectors per batch (0 = skip)")
    ap.add_argument("--c5-launch", type=int, default=256, help="large-workload vectors per pipeline launch")
    ap.add_argument("--c5-steps", type=int, default=1)
    ap.add_argument("--c5-check", type=int, default=256,
                    help="large-workload draws re-run in FP64 (their finite ones give |dlnL|)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="MiB copied device to device once after the timed region: a plain copy of "
                         "known size that calibrates the HBM byte counters of a profiling run")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the likelihood path would shard requests across "
              "GPUs with no exchange between them; run it with --gpus 1", file=sys.stderr)
        sys.exit(2)

    # the CPU baseline first: its process pool forks before the GPU is touched
    cpu = None if args.no_cpu else cpu_baseline(args)
    log("cpu baseline done")

    import torch
    from xpsi_amd import configs
    from xpsi_amd.pipeline import LikelihoodPipeline, params_from_model

    torch.cuda.init()
    if args.fold_ec:
        from xpsi_amd import lib
        lib().xh_set_fold_energy_chunk(args.fold_ec)
    if args.wave_chain >= 0:
        from xpsi_amd import lib
        lib().xh_set_wave_fold_chain(args.wave_chain)
    dev = torch.cuda.get_device_name(0)
    stream = torch.cuda.current_stream().cuda_stream

    # ---------------- plain FP64 GPU path, one request at a time (config 1)
    m1 = configs.config1()
    p1 = params_from_model(m1)
    pl1 = LikelihoodPipeline(m1, max_batch=1)
    for _ in range(3):
        pl1.evaluate([p1], stream)
    lat = []
    for _ in range(args.single):
        t0 = time.perf_counter()
        l1 = pl1.evaluate([p1], stream)[0]
        lat.append(time.perf_counter() - t0)
    lat = np.array(lat) * 1e3
    single_rate = 1e3 / lat.mean()
    log("plain FP64 single request: %.3f ms" % lat.mean())

    # ---------------- batched form
    draws = configs.config3_draws(args.n)
    acc = [d for d in draws if configs.prior_accepts(d["mass"], d["radius"], 300.0)]
    groups = {}
    for d in acc:
        groups.setdefault(d["num_energies"], []).append(d)

    # fold kernel and tile shape per energy group, measured on the group (untimed)
    from xpsi_amd.pipeline import autotune_fold
    tuned, pls = {}, {}
    for ne, g in sorted(groups.items()):
        cands = None if args.fold == "auto" else ((args.fold.split("/")[0], int(args.fold.split("/")[1])),)
        mdl, tuned[ne] = autotune_fold(dict(configs.config3_model(ne), precision=args.precision), g, stream,
                                       **({} if cands is None else {"candidates": cands}))
        pls[ne] = LikelihoodPipeline(mdl, max_batch=len(g))
        if not args.no_graphs:
            pls[ne].use_graphs(True)   # each group's step replayed as one captured hipGraph

    def make(prec):
        return {ne: LikelihoodPipeline(dict(configs.config3_model(ne), precision=prec), max_batch=len(g))
                for ne, g in sorted(groups.items())}

    res = {}
    # one stream per energy group (default priority), largest group first: the
    # groups' kernels overlap, so a fold launch's tail of uneven members runs
    # beside the next group's mesh and rays
    order = sorted(groups, reverse=True)
    gstreams = {ne: torch.cuda.Stream() for ne in groups}

    def step(multi=not args.one_stream, out=res):
        for ne in order:
            pls[ne].launch(groups[ne], gstreams[ne].cuda_stream if multi else stream)
        for ne in order:
            out[ne] = pls[ne].wait()

    log("autotuned fold shapes: %s" % {ne: t["chosen"] for ne, t in tuned.items()})
    log("fold ms per group and candidate: %s" % {ne: {k: round(v, 2) for k, v in t.items() if k != "chosen"}
                                                  for ne, t in tuned.items()})
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    log("batched form: %.1f ms per step" % (dt / args.steps * 1e3))
    # stage times of the overlapped step: one more step, launched eagerly with
    # HIP events around each stage (outside the timed region)
    for pl in pls.values():
        pl.set_timing(True)
    step(out={})
    torch.cuda.synchronize()
    stage = {k: 0.0 for k in LikelihoodPipeline.STAGES}
    for pl in pls.values():
        for k, v in pl.stage_ms().items():
            stage[k] += v
    # work of the dominant kernel (profile and phase fold), per step
    evals = 0.0
    for ne, g in sorted(groups.items()):
        w = pls[ne].work(len(g))
        evals += float(w.sum()) * ne * pls[ne].N_P
    # the same step with every group on one stream, outside the timed region:
    # its HIP events time each kernel alone, where the overlapped step's
    # events also count the time a kernel shares the GPU with other groups'
    for pl in pls.values():
        pl.set_timing(True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    step(multi=False, out={})   # the timed step's results stay in res
    torch.cuda.synchronize()
    one_stream_ms = (time.perf_counter() - t1) * 1e3
    stage1 = {k: 0.0 for k in LikelihoodPipeline.STAGES}
    for pl in pls.values():
        for k, v in pl.stage_ms().items():
            stage1[k] += v
    flops = 8.0 * evals
    achieved = flops / (stage1["profile_fold"] * 1e-3) / 1e12
    # measured bytes per fold launch (mean of the step's launches), for the
    # default workload and shape only
    traffic = None
    if args.n == 4096 and args.precision == "mixed":
        per = [FOLD_TRAFFIC_KIB.get(tuned[ne]["chosen"], {}).get(ne) for ne in sorted(groups)]
        if all(v is not None for v in per):
            traffic = sum(2.0 * f + w for f, w in per) * 1024.0 / len(per)
    achieved_overlapped = flops / (stage["profile_fold"] * 1e-3) / 1e12
    n_eval = len(acc)
    value = n_eval * args.steps / dt
    lnl_all = np.concatenate([res[ne] for ne in sorted(groups)])
    finite = lnl_all > -1e89

    # accuracy of the measured form against FP64 on the same vectors (untimed)
    if args.precision != "fp64":
        for pl in pls.values():
            pl.close()
        ref64 = make("fp64")
        l64 = np.concatenate([ref64[ne].evaluate(groups[ne], stream) for ne in sorted(groups)])
        for pl in ref64.values():
            pl.close()
        both = finite & (l64 > -1e89)
        d = np.abs(lnl_all[both] - l64[both])
        dlnl = {"max": float(d.max()), "p50": float(np.percentile(d, 50)),
                "p99": float(np.percentile(d, 99)), "n": int(both.sum()),
                "same_llzero_class": bool(np.all(finite == (l64 > -1e89))),
                "within_1e-3": bool(d.max() < 1e-3)}
    else:
        dlnl = {"max": 0.0, "note": "FP64 form measured"}

    # the same vectors alone on single-request pipelines: identical results
    alone_dev = 0.0
    for ne, g in sorted(groups.items()):
        pl_a = LikelihoodPipeline(dict(configs.config3_model(ne), precision=args.precision), max_batch=1)
        for k in range(0, len(g), max(1, len(g) // 4)):
            la = pl_a.evaluate([g[k]], stream)[0]
            if la > -1e89 and res[ne][k] > -1e89:
                alone_dev = max(alone_dev, abs(la - res[ne][k]))
        pl_a.close()

    # batched config-1 requests: the single-request workload, batched, in
    # four pipelines on four streams (default priority) so that one part's
    # mesh and rays overlap another's fold
    nb1 = args.config1_batch
    sizes = [len(a) for a in np.array_split(np.arange(nb1), min(4, nb1))]
    plbs = [LikelihoodPipeline(dict(m1, precision=args.precision), max_batch=k) for k in sizes]
    bstreams = [torch.cuda.Stream() for _ in sizes]

    def step1():
        for pl, k, s in zip(plbs, sizes, bstreams):
            pl.launch([p1] * k, s.cuda_stream)
        return np.concatenate([pl.wait() for pl in plbs])

    step1()
    torch.cuda.synchronize()
    steps1 = max(10, args.steps)      # a step is about 20 ms: time enough of them
    t0 = time.perf_counter()
    for _ in range(steps1):
        lb = step1()
    torch.cuda.synchronize()
    batched1_rate = nb1 * steps1 / (time.perf_counter() - t0)
    log("batched config-1 requests: %.0f lnL/s" % batched1_rate)
    batched1_dlnl = float(np.max(np.abs(lb - l1)))
    for pl in plbs:
        pl.close()

    prec_desc = {"fp64": "f64",
                 "mixed": "f32"}[args.precision]
    out = {
        "metric": "batched lnL evals/s + p50/p99 latency vs plain FP64 GPU, at |dlnL|<1e-3",
        "value": value, "unit": "lnL evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3,
        "streams": ("one" if args.one_stream else "one per energy group, largest first"),
        "graphs": not args.no_graphs,
        "one_stream_ms_per_step": one_stream_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": prec_desc, "data": "synthetic",
        "precision": {"fp64": "FP64 throughout",
                      "mixed": "FP32 intensity + spline fold (f32 MFMA moment form), FP32 MFMA "
                               "response fold with FP64 sums of 4-deep slices, FP64 "
                               "rays/mesh/phase-bin sums/likelihood"}[
                          args.precision],
        "config": {"workload": "batched form: %d prior draws (%d accepted), uneven "
                               "sqrt_num_cells {24..96} x num_energies {64,128,256}, BB spot, "
                               "100 ch x 32 bins" % (args.n, n_eval),
                   "global_batch": n_eval, "parallelism": "none (1 GPU)"},
        "dlnl_vs_fp64": dlnl,
        "plain_fp64_gpu_single": {"evals_per_s": single_rate, "p50_ms": float(np.percentile(lat, 50)),
                                  "p99_ms": float(np.percentile(lat, 99)),
                                  "workload": "config 1 request, batch 1", "lnl": float(l1)},
        "speedup_vs_plain_fp64_gpu": value / single_rate,
        "batched_config1": {"evals_per_s": batched1_rate, "batch": nb1, "streams": len(sizes), "steps": steps1,
                            "speedup_vs_plain_fp64_gpu_same_workload": batched1_rate / single_rate,
                            "max_abs_dlnl_vs_fp64_single": batched1_dlnl},
        "batch_latency_ms": {"p50": dt / args.steps * 1e3},
        "max_abs_dlnl_alone_vs_batched": alone_dev,
        "finite_fraction": float(finite.mean()),
        "stage_ms_per_step": stage,
        "stage_ms_one_stream": stage1,
        "roofline": {"bound": "mfma", "achieved": achieved,
                     "peak": FP32_MFMA_PEAK_TFLOPS if args.precision == "mixed" else FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "traffic": traffic,
                     "kernel": "+".join(sorted({KERNEL_NAMES[t["chosen"].split("/")[0]] for t in tuned.values()})),
                     "evals_per_step": evals,
                     "flop_per_eval": 8,
                     "measured_on": "one-stream pass of the timed step (HIP events on the fold's stream)",
                     "traffic_unit": "bytes per launch, mean over the step's launches (rocprofv3 FETCH_SIZE x 2 + "
                                     "WRITE_SIZE, scaled by a 512 MiB copy; the wave fold's writes are its FP64 "
                                     "atomic flux additions, the block fold's mostly register spills)",
                     "achieved_overlapped": achieved_overlapped},
        "autotune_fold_tiles": tuned,
        "device": dev,
    }
    out["roofline"]["frac"] = out["roofline"]["achieved"] / out["roofline"]["peak"]
    if args.copy_probe > 0:
        src = torch.ones(args.copy_probe << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_probe_bytes"] = src.numel()
        del src, dst
    if args.c5_n > 0:
        out["config5"] = config5_section(args, stream)
        log("large workload done")
    if not args.no_serving:
        out["serving"] = serving(args, stream_free=True)
        log("serving done")
    if cpu is not None:
        out["cpu_baseline"] = cpu
        out["speedup_vs_cpu_baseline"] = value / cpu["value"]
    print(json.dumps(out))


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_worker(seconds):
    """One request per core: config-1 evaluations on one thread for `seconds`."""
    from xpsi_amd import configs
    from oracle import oracle as O
    O.select_build("fast")
    O.set_threads(1)
    m1 = configs.config1()
    O.loglikelihood(m1)
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        O.loglikelihood(m1)
        n += 1
    return n, time.perf_counter() - t0


def cpu_baseline(args):
    """The C oracle built with the reference's flags (-O3 -march=native
    -funroll-loops -fopenmp, xpsi setup.py:118-126; OpenMP over rings as the
    reference's prange) on the host's cores, config-1 request, three legs:
    one thread (the reference's production setting, Likelihood.py:43-50),
    OpenMP over all cores, and one request per core in a process pool (the
    reference's one-MPI-process-per-core sampling)."""
    import multiprocessing as mp
    from xpsi_amd import configs
    from oracle import oracle as O
    affinity = len(os.sched_getaffinity(0))
    cores = max(1, min(16, affinity))          # the GPU box's CPU share is 16
    budget = args.cpu_seconds
    # one request per core: fresh interpreters (spawned before this process
    # touches the GPU), one BLAS/OpenMP thread each
    ctx = mp.get_context("spawn")
    saved = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS")}
    os.environ.update(OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1")
    try:
        log("cpu: %d worker processes for %.1f s" % (cores, budget / 3.0))
        with ctx.Pool(cores) as pool:
            res = pool.map(_cpu_worker, [budget / 3.0] * cores)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    per_core_rate = sum(n / dt for n, dt in res)
    log("cpu: one request per core on %d processes: %.1f lnL/s" % (cores, per_core_rate))
    O.select_build("fast")
    m1 = configs.config1()

    def lat_leg(threads, nmin, seconds):
        O.set_threads(threads)
        O.loglikelihood(m1)
        lat, t0 = [], time.perf_counter()
        while len(lat) < nmin or time.perf_counter() - t0 < seconds:
            t1 = time.perf_counter()
            O.loglikelihood(m1)
            lat.append(time.perf_counter() - t1)
            if len(lat) >= nmin and time.perf_counter() - t0 > seconds:
                break
        lat = np.array(lat) * 1e3
        return {"threads": threads, "n": len(lat), "p50_ms": float(np.percentile(lat, 50)),
                "p99_ms": float(np.percentile(lat, 99)), "evals_per_s": 1e3 / float(lat.mean())}

    one = lat_leg(1, 30, budget / 3.0)
    log("cpu: one thread %.1f ms per evaluation" % one["p50_ms"])
    allc = lat_leg(cores, 30, budget / 6.0)
    log("cpu: %d OpenMP threads %.1f ms per evaluation" % (cores, allc["p50_ms"]))
    O.set_threads(1)
    return {"value": one["evals_per_s"], "unit": "lnL evals/s", "cores": 1, "kind": "port",
            "sample": "config-1 requests one at a time on one thread (%d evaluations), the C "
                      "oracle built -O3 -march=native -funroll-loops -fopenmp" % one["n"],
            "single_thread": one, "all_cores_openmp": allc,
            "one_request_per_core": {"processes": cores, "evals_per_s": per_core_rate,
                                     "seconds": budget / 3.0},
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(), "affinity": affinity}


def config5_section(args, stream):
    """The large workload: two hot regions (spot + ceding ring each) with an
    elsewhere component and interstellar absorption, sqrt_num_cells 256, 256
    leaves, 512 energies, 64 phase bins; prior draws evaluated in one launch
    in the measured precision, and the first few again in FP64 for |dlnL|."""
    import torch
    from xpsi_amd import configs
    from xpsi_amd.pipeline import LikelihoodPipeline
    model = configs.config5("full")
    draws = [d for d in configs.config5_draws(4 * args.c5_n)
             if configs.prior_accepts(d["mass"], d["radius"], 300.0)][:args.c5_n]
    # one batch of args.c5_n vectors, run as launches of args.c5_launch on one
    # pipeline (its buffers sized for the launch)
    pl = LikelihoodPipeline(dict(model, precision=args.precision), max_batch=min(len(draws), args.c5_launch))
    pl.evaluate(draws[:pl.max_batch], stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = max(1, args.c5_steps)
    for _ in range(reps):
        got = pl.evaluate(draws, stream)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    pl_launch = pl.max_batch
    pl.close()
    out = {"evals_per_s": len(draws) / dt, "batch": len(draws), "vectors_per_launch": pl_launch,
           "ms_per_batch": dt * 1e3,
           "finite_fraction": float(np.mean(got > -1e89)),
           "workload": "2 hot regions (super + cede) + elsewhere (64x64 cells, 1000 rays) + ISM, "
                       "sqrt_num_cells 256, 256 leaves, 512 energies, 300 ch x 64 bins, %s form"
                       % args.precision}
    nchk = min(args.c5_check, len(draws))
    if args.precision != "fp64" and nchk > 0:
        pl64 = LikelihoodPipeline(dict(model, precision="fp64"), max_batch=min(nchk, args.c5_launch))
        l64 = pl64.evaluate(draws[:nchk], stream)
        pl64.close()
        ok = (l64 > -1e89) & (got[:nchk] > -1e89)
        d = np.abs(got[:nchk][ok] - l64[ok])
        out["dlnl_vs_fp64"] = {"n": int(ok.sum()), "max": float(d.max()) if d.size else None,
                               "p50": float(np.percentile(d, 50)) if d.size else None,
                               "p99": float(np.percentile(d, 99)) if d.size else None,
                               "within_1e-3": bool(d.size and d.max() < 1e-3),
                               "same_llzero_class": bool(np.all((l64 > -1e89) == (got[:nchk] > -1e89)))}
    return out


def serving(args, stream_free=True):
    """The serving form on its workload: a stream of config-2-model requests
    (two regions with ceding rings, 4-D table atmosphere, 256 energies, 300
    channels, 128 leaves) with mixed sqrt_num_cells {32..96}, into the
    continuous-batching server (batches of 1..slot_batch formed under a
    latency cap, admitted by device footprint, graphs captured for every batch
    size beforehand).  Reported: the sustained maximum (a closed-loop burst),
    then Poisson arrivals offered at 25%, 50%, 80% and 95% of that maximum with
    p50/p99 submit-to-result latency; the plain FP64 single-request loop of
    the same requests beside."""
    from xpsi_amd import configs
    from xpsi_amd.pipeline import LikelihoodPipeline
    from xpsi_amd.serve import LikelihoodServer
    model = configs.config4_model()
    draws = [d for d in configs.config4_draws(4096)
             if configs.prior_accepts(d["mass"], d["radius"], 300.0)]
    # the plain FP64 single-request loop on the same requests
    pl = LikelihoodPipeline(model, max_batch=1)
    for d in draws[:3]:
        pl.evaluate([d])
    lat1 = []
    for d in draws[:args.single]:
        t0 = time.perf_counter()
        pl.evaluate([d])
        lat1.append(time.perf_counter() - t0)
    pl.close()
    lat1 = np.array(lat1) * 1e3
    srv = LikelihoodServer(dict(model, precision=args.precision), n_slots=args.serve_slots,
                           slot_batch=args.serve_batch, hbm_budget_bytes=args.serve_budget,
                           use_graphs=True, max_wait_s=args.serve_wait)
    log("serving: plain single requests %.3f ms; capturing graphs" % float(lat1.mean()))
    t0 = time.perf_counter()
    srv.warm()
    warm_s = time.perf_counter() - t0
    log("serving: graphs captured in %.1f s" % warm_s)

    def drain(n, timeout=120.0):
        lat, done, t0 = [], 0, time.perf_counter()
        while done < n and time.perf_counter() - t0 < timeout:
            ids, lnl, l = srv.poll()
            done += len(ids)
            lat.extend(l.tolist())
        return done, lat

    # sustained maximum: a burst of requests submitted at once
    nb = min(len(draws), 2 * args.serve_batch * args.serve_slots)
    t0 = time.perf_counter()
    for d in draws[:nb]:
        srv.submit(d)
    done, _ = drain(nb)
    sustained = done / (time.perf_counter() - t0)
    out = {"sustained_max_req_per_s": sustained, "burst": nb}
    log("serving: sustained %.0f req/s" % sustained)
    rng = np.random.default_rng(4)
    for load in (0.25, 0.5, 0.8, 0.95):
        rate = load * sustained
        n = max(64, int(rate * args.serve_seconds))
        t_sub = np.cumsum(rng.exponential(1.0 / rate, n))
        lat, done = [], 0
        t0 = time.perf_counter()
        k = 0
        while done < n:
            now = time.perf_counter() - t0
            while k < n and t_sub[k] <= now:
                srv.submit(draws[k % len(draws)])
                k += 1
            ids, lnl, l = srv.poll()
            done += len(ids)
            lat.extend(l.tolist())
            if time.perf_counter() - t0 > 60:
                break
        wall = time.perf_counter() - t0
        lat = np.array(lat) * 1e3
        out["load_%.2f" % load] = {
            "offered_req_per_s": rate, "submitted": int(k), "completed": int(done), "throughput": done / wall,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99))}
        log("serving at %.0f req/s (%.0f%% load): p50 %.2f ms p99 %.2f ms"
            % (rate, 100 * load, np.percentile(lat, 50), np.percentile(lat, 99)))
        if done < n:
            drain(k - done)
    st = srv.stats()
    srv.close()
    out["stats"] = st
    out["graph_warm_s"] = warm_s
    out["plain_fp64_single_same_requests"] = {
        "evals_per_s": 1e3 / float(lat1.mean()), "p50_ms": float(np.percentile(lat1, 50)),
        "p99_ms": float(np.percentile(lat1, 99)), "n": int(len(lat1))}
    out["speedup_sustained_vs_plain_fp64_single"] = sustained * float(lat1.mean()) / 1e3
    out["workload"] = ("config-2 model requests (2 regions + ceding rings, 4-D table, 256 energies, "
                       "300 ch), sqrt_num_cells {32,48,64,96}; %d slots x %d, latency cap %.1f ms, "
                       "HBM budget %.0f GB, graphs on, %s form"
                       % (args.serve_slots, args.serve_batch, args.serve_wait * 1e3,
                          args.serve_budget / 1e9, args.precision))
    return out


if __name__ == "__main__":
    main()
