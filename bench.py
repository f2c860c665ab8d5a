#!/usr/bin/env python3
"""Benchmark of the correlation hot path on one MI355X.

One step = one pass of the pair kernel over every forest pair of the synthetic
100k-forest survey (all healpix rows in one launch), with the forest store and the pair
list already resident in HBM.  The metric is weighted pixel pairs per second: the number
of (i, j) pixel pairs with w1 != 0 and w2 != 0 over all neighbouring forest pairs -- the
iterations of the reference's inner loop (py/picca/cf.py:330-387) -- divided by the wall
time of one step.

Beside it:
  * roofline of the pair kernel (FP64 vector ALU): the reference algorithm's FP64
    operations for this launch (4 per evaluated pixel pair, 17 more per accepted one) over
    the kernel's duration, taken with HIP events on the stream the kernel runs on;
  * a CPU baseline: the FP64 oracle (a scalar C restatement of the reference loop) on a
    seeded random sample of the same forest pairs, on this host's cores;
  * parity on that sample: num_pairs mismatches against the oracle and the largest
    relative xi error.

    python bench.py [--steps K] [--warmup W] [--workload survey_100k] [--scale 1.0]
                    [--copy-probe MIB]

--copy-probe runs a plain device-to-device copy of MIB mebibytes before the steps, so that
a hardware-counter profile of the same command holds a kernel of known traffic to scale
FETCH_SIZE / WRITE_SIZE by.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "weighted pixel pairs/sec (1 GPU) vs numba host; xi err vs FP64; fits/s"
FP64_VALU_PEAK_TFLOPS = 78.6   # MI355X FP64 vector and dense FP64 matrix, published
HBM_PEAK_GBS = 8000.0
# FP64 operations per pair in the reference's loop (r_par sub+mul, r_trans add+mul for
# every evaluated pair; z 2, bins 5, products 3, accumulations 5 adds + 3 muls more for
# an accepted one)
FLOPS_EVALUATED = 4
FLOPS_ACCEPTED_EXTRA = 17


def log(msg):
    sys.stderr.write("[bench] %s\n" % msg)
    sys.stderr.flush()


class Globals:
    num_bins_r_par = 50
    num_bins_r_trans = 50
    num_model_bins_r_par = 0
    num_model_bins_r_trans = 0
    r_par_min = 0.0
    r_par_max = 200.0
    r_trans_max = 200.0
    z_min_pairs = None
    z_max_pairs = None
    zerr_cut_deg = None
    zerr_cut_kms = None
    z_ref = 2.25
    alpha = 2.9
    alpha2 = 2.9
    x_correlation = False
    rmu_binning = False
    ang_correlation = False
    remove_same_half_plate_close_pairs = False
    redshift_evolution_in_distortion_matrix = True


def cpu_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def measured_traffic(workload, scale, kernel="xi_kernel"):
    """HBM bytes of one step of the pair kernel (all its launches), measured with FETCH_SIZE +
    WRITE_SIZE (each scaled by the same counters on a plain copy), as recorded beside the
    profiles."""
    path = os.path.join(ROOT, "profiles", "%s_hbm_traffic.json" % kernel)
    try:
        with open(path) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return 0.0
    entry = table.get("%s@%g" % (workload, scale))
    return float(entry["traffic_bytes"]) if entry else 0.0


def copy_probe(eng, mib):
    """A plain device copy kernel over `mib` MiB (calibrates HBM byte counters)."""
    nbytes = int(mib * (1 << 20)) // 16 * 16
    ms = eng.diag_copy(nbytes)
    log("copy probe: %d bytes read, %d bytes written, %.3f ms (%.0f GB/s)"
        % (nbytes, nbytes, ms, 2 * nbytes / (ms / 1e3) / 1e9))


class StepRunner:
    """Warm-up and timed steps under a wall-clock budget.

    Runs up to `warmup` untimed steps (stopping early once they have used a fifth of the
    budget) and then up to `steps` timed steps, stopping before a step that would end past
    the budget; at least one timed step always runs.  `warmup_done` and `timings` hold what
    actually ran, which is what the JSON line reports.
    """

    def __init__(self, warmup, steps, budget_s):
        self.warmup = max(int(warmup), 0)
        self.steps = max(int(steps), 1)
        self.budget_s = float(budget_s)
        self.warmup_done = 0
        self.timings = []
        self.results = []

    def run(self, step, label="step"):
        start = time.perf_counter()
        last = 0.0
        for w in range(self.warmup):
            ts = time.perf_counter()
            step()
            last = time.perf_counter() - ts
            self.warmup_done += 1
            log("warmup %d: %.1f ms" % (w, last * 1e3))
            if time.perf_counter() - start >= 0.2 * self.budget_s:
                break
        for k in range(self.steps):
            elapsed = time.perf_counter() - start
            expect = float(np.mean(self.timings)) if self.timings else last
            if k > 0 and elapsed + expect > self.budget_s:
                log("time budget of %.0f s reached after %d timed steps" % (self.budget_s, k))
                break
            ts = time.perf_counter()
            out = step()
            self.timings.append(time.perf_counter() - ts)
            self.results.append(out)
            log("%s %d: wall %.1f ms" % (label, k, self.timings[-1] * 1e3))
        return self

    @property
    def steps_done(self):
        return len(self.timings)

    @property
    def mean_s(self):
        return float(np.mean(self.timings))

    def record(self):
        """The requested counts and the budget, for the JSON line's config."""
        return {"warmup_requested": self.warmup, "steps_requested": self.steps,
                "time_budget_s": self.budget_s}


def sample_pairs(pairs, store, target_evaluated, seed=1):
    """Seeded random subset of forest pairs holding about target_evaluated pixel pairs."""
    nz = store.nonzero_weight_counts()
    per_pair = nz[pairs.pair_a].astype(np.float64) * nz[pairs.pair_b]
    mean = max(per_pair.mean(), 1.0)
    count = int(min(pairs.num_pairs, max(16, target_evaluated / mean)))
    rng = np.random.default_rng(seed)
    sel = np.sort(rng.choice(pairs.num_pairs, count, replace=False))
    from picca_amd.batch import XiPairs
    return XiPairs(pairs.pair_a[sel], pairs.pair_b[sel], pairs.pair_ang[sel],
                   pairs.pair_flags[sel], pairs.pair_row[sel], pairs.healpixs)


def main():
    parser = argparse.ArgumentParser(description=__doc__,
                                     formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=2)
    parser.add_argument("--warmup", type=int, default=1)
    parser.add_argument("--workload", default="survey_100k")
    parser.add_argument("--scale", type=float, default=1.0,
                        help="fraction of the sky area (and forests) of the workload")
    parser.add_argument("--seed", type=int, default=0)
    parser.add_argument("--cpu-pairs", type=float, default=1.2e10,
                        help="evaluated pixel pairs in the CPU-baseline sample")
    parser.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    parser.add_argument("--time-budget", type=float, default=120.0,
                        help="wall seconds for the warm-up and timed GPU steps together; "
                             "the steps that fit are run and reported")
    parser.add_argument("--copy-probe", type=float, default=0.0,
                        help="MiB of a plain device copy to run before the steps")
    parser.add_argument("--traffic-bytes", type=float, default=-1.0,
                        help="HBM bytes per launch of the pair kernel (default: the "
                             "counter-derived figure in profiles/ for this workload)")
    parser.add_argument("--mode", default="xi",
                        choices=("xi", "dmat", "continuum", "precision", "xcf", "xdmat",
                                 "metal", "cov", "wick"),
                        help="xi: the pair kernel (default); dmat: the distortion matrix; "
                             "continuum: plain vs learned-seed continuum fits; precision: "
                             "xi error and edge pairs on the heavy-tailed survey; xcf / xdmat: "
                             "forest x quasar correlation / distortion matrix; metal: the "
                             "LYA x SiIII(1207) metal distortion matrix; cov: the "
                             "subsampling covariance of picca_export; wick: the Wick "
                             "covariance diagrams 1-3")
    parser.add_argument("--reject", type=float, default=0.99,
                        help="dmat: fraction of neighbour pairs rejected")
    parser.add_argument("--wick-pairs", type=int, default=64,
                        help="wick: forest pairs of the workload")
    parser.add_argument("--wick-cpu-pairs", type=int, default=3,
                        help="wick: forest pairs of the CPU-baseline sample")
    parser.add_argument("--schedule", default="auto", choices=("auto", "pairs", "runs"),
                        help="xi: schedule of the pair kernel (auto: the pair loop, measured "
                             "faster on these grids; runs: the run walk where it applies)")
    parser.add_argument("--spectra", type=int, default=200000,
                        help="continuum: spectra fitted (and as many more for training)")
    args = parser.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: this engine has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.mode == "dmat":
        return bench_dmat(args)
    if args.mode == "continuum":
        return bench_continuum(args)
    if args.mode == "precision":
        return bench_precision(args)
    if args.mode in ("xcf", "xdmat"):
        return bench_xcf(args)
    if args.mode == "metal":
        return bench_metal(args)
    if args.mode == "cov":
        return bench_cov(args)
    if args.mode == "wick":
        return bench_wick(args)
    return bench_xi(args)


def bench_xi(args):
    from picca_amd import batch, synth
    from picca_amd.constants import Cosmo
    from picca_amd.engine import device_info, get_engine, params_from_globals
    from picca_amd.utils import compute_ang_max

    t0 = time.time()
    survey = synth.make_survey(args.workload, seed=args.seed, scale=args.scale)
    store = survey.store
    log("workload %s scale %g: %d forests, %d pixels (%.1f s)"
        % (args.workload, args.scale, store.num_forests, store.num_pixels, time.time() - t0))
    ang_max = compute_ang_max(Cosmo(0.315, verbose=False), Globals.r_trans_max, survey.z_min)
    t1 = time.time()
    pairs = batch.xi_pairs(store, survey.healpix, ang_max)
    neighbour_s = time.time() - t1
    evaluated = pairs.evaluated_pixel_pairs(store)
    log("neighbours: %d forest pairs in %d rows, %.3e evaluated pixel pairs (%.1f s)"
        % (pairs.num_pairs, pairs.num_rows, evaluated, neighbour_s))

    params = params_from_globals(Globals)
    eng = get_engine(0)
    info = device_info(0)
    from picca_amd import _native
    eng.set_xi_schedule({"auto": _native.XI_SCHEDULE_AUTO, "pairs": _native.XI_SCHEDULE_PAIRS,
                         "runs": _native.XI_SCHEDULE_RUNS}[args.schedule])
    t2 = time.time()
    eng.set_forests(store, key="bench")
    eng.set_xi_pairs(pairs.pair_a, pairs.pair_b, pairs.pair_ang, pairs.pair_flags,
                     pairs.pair_row, pairs.num_rows, key="bench")
    upload_s = time.time() - t2
    log("device %s: upload %.1f s" % (info["name"], upload_s))

    if args.traffic_bytes < 0:
        args.traffic_bytes = measured_traffic(args.workload, args.scale)
    if args.copy_probe > 0:
        copy_probe(eng, args.copy_probe)

    # each step returns after the stream has drained
    runner = StepRunner(args.warmup, args.steps, args.time_budget).run(
        lambda: eng.xi_execute(params))
    report = runner.results[-1]
    ms_per_step = 1e3 * runner.mean_s
    kernel_ms = float(np.mean([r.kernel_ms for r in runner.results]))
    log("pair kernel: %.1f ms per step (HIP events)" % kernel_ms)
    value = evaluated / (ms_per_step / 1e3)
    rep = report.as_dict()

    # roofline of the pair kernel: the reference loop's FP64 operations for this launch
    # (every evaluated pair, pruned or not, and the accepted ones), over the kernel time
    flops = FLOPS_EVALUATED * evaluated + FLOPS_ACCEPTED_EXTRA * rep["accepted_pairs"]
    achieved = flops / (kernel_ms / 1e3) / 1e12
    # the operations the kernel itself executes: it walks only the pixel pairs left after
    # the r_par / r_trans window pruning
    executed = FLOPS_EVALUATED * rep["visited_pairs"] + FLOPS_ACCEPTED_EXTRA * rep["accepted_pairs"]
    roofline = {"bound": "valu", "achieved": achieved, "peak": FP64_VALU_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP64_VALU_PEAK_TFLOPS,
                "traffic": args.traffic_bytes if args.traffic_bytes > 0 else None,
                "flops_per_launch": flops,
                "executed_tflops": executed / (kernel_ms / 1e3) / 1e12,
                # the run-walk instance runs whenever any task took the walk
                "kernel": "picca::xi_kernel<0,false,false,true,%s>"
                          % ("true" if rep["walk_tasks"] > 0 else "false"),
                "kernel_ms": kernel_ms,
                "hbm_gbs_at_traffic": (args.traffic_bytes / (kernel_ms / 1e3) / 1e9
                                       if args.traffic_bytes > 0 else None)}

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "pixel pairs/s",
        "n_gpus": 1,
        "steps": runner.steps_done,
        "warmup": runner.warmup_done,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "scale": args.scale,
            "forests": store.num_forests,
            "pixels": store.num_pixels,
            "forest_pairs": pairs.num_pairs,
            "healpix_rows": pairs.num_rows,
            "evaluated_pixel_pairs": evaluated,
            "accepted_pixel_pairs": rep["accepted_pairs"],
            "kernel_visited_pixel_pairs": rep["visited_pairs"],
            "kernel_wave_steps": rep["wave_steps"],
            "kernel_tasks": rep["tasks"],
            "kernel_empty_tasks": rep["empty_tasks"],
            "schedule": args.schedule,
            # what the kernel ran: the pair loop over 64-pixel chunks of delta1 (one wave
            # per chunk and neighbour forest), the run walk, or both
            "schedule_taken": ("pairs" if rep["walk_tasks"] == 0 else
                               "runs" if rep["walk_tasks"] == rep["tasks"] - rep["empty_tasks"]
                               else "runs+pairs"),
            "kernel_walk_tasks": rep["walk_tasks"],
            "kernel_runs": rep["runs"],
            "kernel_probes": rep["probes"],
            "edge_pairs": rep["edge_pairs"],
            "overflow_pairs": rep["overflow_pairs"],
            "bins": "50x50", "r_par_max": 200.0, "r_trans_max": 200.0, "nside": survey.nside,
            "z_min": survey.z_min, "ang_max_deg": float(np.degrees(ang_max)),
            "neighbour_search_s": neighbour_s, "upload_s": upload_s,
            "device": info["name"],
            **runner.record(),
        },
        "roofline": roofline,
        "cpu_baseline": None,
    }

    if not args.no_cpu:
        try:
            cpu_leg_xi(args, eng, params, store, pairs, value, result)
        except Exception as exc:  # the JSON line is printed whatever the CPU legs do
            log("cpu baseline failed: %r" % (exc,))
            result["cpu_baseline_error"] = repr(exc)
    print(json.dumps(result))


def cpu_leg_xi(args, eng, params, store, pairs, value, result):
    """CPU baseline and parity sample of the xi bench.

    The FP64 C oracle runs on a seeded random sample of the forest pairs, one OpenMP task
    per healpix row, on every core this process may use; a second, smaller sample runs with
    picca_cf's default process count, cpu_count() // 2 (py/picca/bin/picca_cf.py:340-341),
    whatever the affinity allows.  The GPU then runs the first sample for parity.
    """
    from oracle import cf_oracle as O
    sample = sample_pairs(pairs, store, args.cpu_pairs)
    sample_eval = sample.evaluated_pixel_pairs(store)
    cores = cpu_cores()
    log("cpu baseline: %d forest pairs, %.3e pixel pairs on %d threads"
        % (sample.num_pairs, sample_eval, cores))
    oparams = O.params_from_module(Globals)
    tc = time.time()
    out, stats = O.xi_rows(oparams, store, sample.pair_a, sample.pair_b, sample.pair_ang,
                           sample.pair_flags, sample.pair_row, sample.num_rows,
                           num_threads=cores)
    cpu_s = time.time() - tc
    cpu_rate = sample_eval / cpu_s
    log("cpu baseline: %.1f s, %.3e pixel pairs/s" % (cpu_s, cpu_rate))
    result["cpu_baseline"] = {
        "value": cpu_rate, "unit": "pixel pairs/s", "cores": cores, "kind": "port",
        "sample": "%d random forest pairs (seed 1) of the same workload, %.3e pixel pairs, "
                  "FP64 C oracle (scalar restatement of the reference loop), one OpenMP "
                  "task per healpix row" % (sample.num_pairs, sample_eval),
        "seconds": cpu_s,
    }
    result["speedup_vs_cpu"] = value / cpu_rate

    # picca_cf's default worker count
    default_threads = max(1, (os.cpu_count() or 2) // 2)
    small = sample_pairs(pairs, store, args.cpu_pairs / 3, seed=2)
    small_eval = small.evaluated_pixel_pairs(store)
    tc = time.time()
    O.xi_rows(oparams, store, small.pair_a, small.pair_b, small.pair_ang, small.pair_flags,
              small.pair_row, small.num_rows, num_threads=default_threads)
    small_s = time.time() - tc
    log("cpu baseline at cpu_count()//2 = %d threads: %.1f s, %.3e pixel pairs/s"
        % (default_threads, small_s, small_eval / small_s))
    result["cpu_baseline"]["picca_default_nproc"] = {
        "value": small_eval / small_s, "threads": default_threads,
        "cores_available": cores,
        "sample": "%d random forest pairs (seed 2), %.3e pixel pairs"
                  % (small.num_pairs, small_eval)}

    # GPU on the first sample: parity with the FP64 oracle
    eng.set_xi_pairs(sample.pair_a, sample.pair_b, sample.pair_ang, sample.pair_flags,
                     sample.pair_row, sample.num_rows, key="sample")
    gres = eng.xi(params)
    nb_mismatch = int(np.count_nonzero(gres.num_pairs != out[5]))
    we_o = out[0]
    mask = we_o > 0
    xi_g = gres.xi[mask] / gres.weights[mask]
    xi_o = out[1][mask] / we_o[mask]
    rel = np.abs(xi_g - xi_o) / np.maximum(np.abs(xi_o), 1e-300)
    result["parity_sample"] = {
        "num_pairs_mismatches": nb_mismatch,
        "accepted_pairs_gpu": int(gres.report["accepted_pairs"]),
        "accepted_pairs_oracle": int(stats[0]),
        "edge_pairs_gpu": int(gres.report["edge_pairs"]),
        "edge_pairs_oracle": int(stats[1]),
        "xi_max_rel_err": float(rel.max()) if rel.size else 0.0,
        "xi_median_rel_err": float(np.median(rel)) if rel.size else 0.0,
    }



class DmatGlobals(Globals):
    num_model_bins_r_par = 100
    num_model_bins_r_trans = 50


class MetalGlobals(Globals):
    num_model_bins_r_par = 50
    num_model_bins_r_trans = 50


class WickGlobals(Globals):
    num_bins_r_par = 15
    num_bins_r_trans = 15
    r_par_max = 60.0
    r_trans_max = 60.0


def bench_dmat(args):
    """Distortion matrix of the 100k-forest survey: neighbour pairs kept with probability
    1 - reject (one seeded legacy-numpy draw per pair, in pair-list order), 50x50 data bins,
    100x50 model bins, redshift evolution on."""
    from picca_amd import batch, synth
    from picca_amd.constants import Cosmo
    from picca_amd.engine import device_info, get_engine, params_from_globals
    from picca_amd.utils import compute_ang_max

    survey = synth.make_survey("survey_100k", seed=args.seed, scale=args.scale)
    store = survey.store
    ang_max = compute_ang_max(Cosmo(0.315, verbose=False), Globals.r_trans_max, survey.z_min)
    pairs = batch.xi_pairs(store, survey.healpix, ang_max)
    keep = np.random.RandomState(args.seed).rand(pairs.num_pairs) > args.reject
    a, b = pairs.pair_a[keep], pairs.pair_b[keep]
    ang, flags = pairs.pair_ang[keep], pairs.pair_flags[keep]
    log("dmat: %d of %d neighbour pairs kept" % (len(a), pairs.num_pairs))
    params = params_from_globals(DmatGlobals)
    eng = get_engine(0)
    info = device_info(0)
    eng.set_forests(store, key="bench_dmat")
    for w in range(args.warmup):
        eng.dmat(params, a, b, ang, flags)
    times = []
    out = None
    for k in range(args.steps):
        t = time.perf_counter()
        out = eng.dmat(params, a, b, ang, flags)
        times.append(time.perf_counter() - t)
        log("dmat step %d: %.2f s" % (k, times[-1]))
    sec = float(np.mean(times))
    stats = out[6]
    rep = eng.dmat_report()
    result = {
        "metric": "distortion-matrix in-range pixel pairs/s (1 GPU)",
        "value": stats[0] / sec, "unit": "pixel pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * sec, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "dmat_100k", "scale": args.scale, "forests": store.num_forests,
                   "neighbour_pairs": pairs.num_pairs, "kept_pairs": int(len(a)),
                   "reject": args.reject, "in_range_pairs": int(stats[0]),
                   "selected_pairs": int(stats[1]), "bins": "50x50", "model_bins": "100x50",
                   "device": info["name"], "timing": "wall time of picca_dmat_run, "
                   "including the 100 MB download of the matrix"},
        # the fill kernel on the FP64 matrix cores: the operations it issues there per launch
        # sequence (16x16x4 products, 2048 FLOP each) over its HIP-event time
        "roofline": {"bound": "mfma", "unit": "TFLOP/s", "peak": FP64_VALU_PEAK_TFLOPS,
                     "achieved": rep["mfma_flops"] / (rep["fill_ms"] * 1e-3) / 1e12,
                     "frac": rep["mfma_flops"] / (rep["fill_ms"] * 1e-3) / 1e12
                     / FP64_VALU_PEAK_TFLOPS, "traffic": None,
                     "kernel": "picca::dmat_fill_kernel<false>", "kernel_ms": rep["fill_ms"],
                     "pass_kernel_ms": rep["pass_ms"], "batches": rep["batches"],
                     "mfma_flops": rep["mfma_flops"]},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        from oracle import cf_oracle as O
        rng = np.random.default_rng(1)
        sel = np.sort(rng.choice(len(a), min(len(a), 12), replace=False))
        oparams = O.params_from_module(DmatGlobals)
        tc = time.time()
        ref, ref_stats = O.dmat_rows(oparams, store, a[sel], b[sel], ang[sel], flags[sel])
        cpu_s = time.time() - tc
        mine = eng.dmat(params, a[sel], b[sel], ang[sel], flags[sel])
        scale = np.abs(ref[1]).max()
        err = float(np.abs(mine[1] - ref[1]).max() / scale)
        result["cpu_baseline"] = {
            "value": ref_stats[0] / cpu_s, "unit": "pixel pairs/s", "cores": 1, "kind": "port",
            "sample": "%d random kept pairs (seed 1), %d in-range pixel pairs, FP64 C oracle "
                      "(scalar restatement of compute_dmat_forest_pairs_fast)"
                      % (len(sel), int(ref_stats[0])), "seconds": cpu_s}
        result["parity_sample"] = {"dmat_max_abs_err_over_max": err,
                                   "counts_equal": bool(list(mine[6]) == list(ref_stats))}
    print(json.dumps(result))


# FP64 operations of one pass of compute_metal_dmat per pixel pair
# (py/picca/cf.py:1026-1090): Lya separations (sub, mul, add, mul) and data bins (sub, div,
# mul, div, mul) for every evaluated pair; absorber separations, model bins and w1 w2 for a
# pair with an in-range data bin; the evolution factor, the five weighted terms and their
# five sums for a pair with an in-range model bin.
METAL_FLOPS_VISITED = 9
METAL_FLOPS_DATA = 10
METAL_FLOPS_MODEL = 15


def bench_metal(args):
    """Metal distortion matrix LYA x SiIII(1207) of the 100k-forest survey: neighbour pairs
    kept with probability 1 - reject (one seeded legacy-numpy draw per pair), 50x50 data
    and model bins, both absorber passes of the auto-correlation."""
    from picca_amd import batch, metal, synth
    from picca_amd import _native
    from picca_amd.constants import ABSORBER_IGM, Cosmo
    from picca_amd.engine import device_info, get_engine, params_from_globals
    from picca_amd.utils import compute_ang_max

    survey = synth.make_survey("survey_100k", seed=args.seed, scale=args.scale)
    store = survey.store
    cosmo = Cosmo(0.315, verbose=False)
    ang_max = compute_ang_max(cosmo, Globals.r_trans_max, survey.z_min)
    pairs = batch.xi_pairs(store, survey.healpix, ang_max)
    keep = np.random.RandomState(args.seed).rand(pairs.num_pairs) > args.reject
    a, b = pairs.pair_a[keep], pairs.pair_b[keep]
    ang, flags = pairs.pair_ang[keep], pairs.pair_flags[keep]
    log("metal: %d of %d neighbour pairs kept" % (len(a), pairs.num_pairs))
    t = time.time()
    lya = metal.absorber_table(store, ABSORBER_IGM["LYA"], 2.9, cosmo)
    si = metal.absorber_table(store, ABSORBER_IGM["SiIII(1207)"], 1.0, cosmo)
    den = (1 + MetalGlobals.z_ref) ** (2.9 + 1.0 - 2)
    log("metal: absorber tables (%.1f s)" % (time.time() - t))
    params = params_from_globals(MetalGlobals)
    mode = _native.PICCA_METAL_AUTO
    eng = get_engine(0)
    info = device_info(0)
    eng.set_forests(store, key="bench_metal")
    eng.metal_set_absorber(0, *lya)
    eng.metal_set_absorber(1, *si)
    if args.traffic_bytes < 0:
        args.traffic_bytes = measured_traffic("metal_dmat_100k", args.scale, "metal_kernel")
    if args.copy_probe > 0:
        copy_probe(eng, args.copy_probe)
    nb = MetalGlobals.num_bins_r_par * MetalGlobals.num_bins_r_trans
    nm = MetalGlobals.num_model_bins_r_par * MetalGlobals.num_model_bins_r_trans

    def step():
        out = metal.outputs(nb, nm)
        reps = [eng.metal_pass(params, mode, 0, 1, den, a, b, ang, flags, out),
                eng.metal_pass(params, mode, 1, 0, den, a, b, ang, flags, out)]
        return out, reps

    for w in range(args.warmup):
        step()
    wall, kernel, reps = [], [], None
    for k in range(args.steps):
        t = time.perf_counter()
        out, reps = step()
        wall.append(time.perf_counter() - t)
        kernel.append(sum(r["kernel_ms"] for r in reps))
        log("metal step %d: wall %.1f ms, kernels %.1f ms" % (k, wall[-1] * 1e3, kernel[-1]))
    kernel_ms = float(np.mean(kernel))
    visited = sum(r["visited_pairs"] for r in reps)
    data_pairs = sum(r["data_pairs"] for r in reps)
    model_pairs = sum(r["model_pairs"] for r in reps)
    flops = (METAL_FLOPS_VISITED * visited + METAL_FLOPS_DATA * data_pairs
             + METAL_FLOPS_MODEL * model_pairs)
    achieved = flops / (kernel_ms / 1e3) / 1e12
    result = {
        "metric": "metal distortion-matrix in-range pixel pairs/s (1 GPU)",
        "value": data_pairs / (kernel_ms / 1e3), "unit": "pixel pairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * float(np.mean(wall)),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "metal_dmat_100k", "scale": args.scale,
                   "forests": store.num_forests, "neighbour_pairs": pairs.num_pairs,
                   "kept_pairs": int(len(a)), "reject": args.reject,
                   "absorbers": "LYA x SiIII(1207), both passes", "bins": "50x50",
                   "model_bins": "50x50", "visited_pixel_pairs": int(visited),
                   "data_pixel_pairs": int(data_pairs), "model_pixel_pairs": int(model_pairs),
                   "device": info["name"],
                   "timing": "value: in-range pixel pairs of both passes over the two "
                             "kernels' HIP-event time; ms_per_step includes the pair upload "
                             "and the 50 MB matrix download of each pass"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_VALU_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_VALU_PEAK_TFLOPS,
                     "traffic": (args.traffic_bytes / 2 if args.traffic_bytes > 0 else None),
                     "flops_per_launch": flops / 2, "launches_per_step": 2,
                     "launch_avg_ms": kernel_ms / 2, "kernel": "picca::metal_kernel",
                     "kernel_ms": kernel_ms},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        from oracle import cf_oracle as O
        rng = np.random.default_rng(1)
        sel = np.sort(rng.choice(len(a), min(len(a), 3000), replace=False))
        oparams = O.params_from_module(MetalGlobals)
        ref = metal.outputs(nb, nm)
        stats = np.zeros(3, dtype=np.int64)
        off = store.offsets
        tc = time.time()
        for k in sel:
            sa = slice(off[a[k]], off[a[k] + 1])
            sb = slice(off[b[k]], off[b[k] + 1])
            for ta, tb in ((lya, si), (si, lya)):
                O.metal_pair(oparams, mode, den, (store.r_comov[sa], store.dist_m[sa]),
                             [x[sa] for x in ta], (store.r_comov[sb], store.dist_m[sb]),
                             [x[sb] for x in tb], ang[k], bool(flags[k]), ref, stats)
        cpu_s = time.time() - tc
        mine = metal.outputs(nb, nm)
        sreps = [eng.metal_pass(params, mode, 0, 1, den, a[sel], b[sel], ang[sel], flags[sel],
                                mine),
                 eng.metal_pass(params, mode, 1, 0, den, a[sel], b[sel], ang[sel], flags[sel],
                                mine)]
        err = max(float(np.abs(m - r).max() / max(np.abs(r).max(), 1e-300))
                  for m, r in zip(mine, ref))
        result["cpu_baseline"] = {
            "value": stats[0] / cpu_s, "unit": "pixel pairs/s", "cores": 1, "kind": "port",
            "sample": "%d random kept pairs (seed 1), both passes, %d in-range pixel pairs, "
                      "FP64 C oracle (scalar restatement of compute_metal_dmat)"
                      % (len(sel), int(stats[0])), "seconds": cpu_s}
        result["speedup_vs_cpu"] = result["value"] / result["cpu_baseline"]["value"]
        result["parity_sample"] = {
            "max_abs_err_over_max": err,
            "counts_equal": bool(sum(r["data_pairs"] for r in sreps) == stats[0]
                                 and sum(r["model_pairs"] for r in sreps) == stats[1])}
    print(json.dumps(result))


def bench_cov(args):
    """Subsampling covariance of picca_export (compute_cov) on COR rows of the 100k-forest
    survey's shape: 1196 healpix rows (nside 64) x 2500 bins (50 x 50), synthetic xi and
    weights; the timed work is the whole computation on inputs resident in HBM."""
    from picca_amd.engine import device_info, get_engine

    rows = max(1, int(round(1196 * args.scale)))
    nb = 2500
    rng = np.random.default_rng(args.seed)
    xi = rng.normal(0.0, 1e-2, size=(rows, nb))
    we = rng.uniform(1e3, 1e5, size=(rows, nb))
    eng = get_engine(0)
    info = device_info(0)
    for w in range(args.warmup):
        eng.cov_subsample(xi, we)
    total, syrk, cov = [], [], None
    for k in range(max(args.steps, 1) * 10):
        cov, ms, ms_syrk = eng.cov_subsample(xi, we)
        total.append(ms)
        syrk.append(ms_syrk)
    total_ms = float(np.mean(total))
    syrk_ms = float(np.mean(syrk))
    flops = float(nb) * (nb + 1) * rows      # upper triangle of Y^T Y, multiply + add
    achieved = flops / (syrk_ms / 1e3) / 1e12
    result = {
        "metric": "subsampling covariances/s (1 GPU)",
        "value": 1e3 / total_ms, "unit": "covariances/s", "n_gpus": 1,
        "steps": len(total), "warmup": args.warmup, "ms_per_step": total_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "cov_100k", "rows": rows, "bins": nb, "device": info["name"],
                   "timing": "HIP events around the two kernels (weighted centring, then "
                             "Y^T Y on the FP64 matrix cores); inputs uploaded before"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_VALU_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_VALU_PEAK_TFLOPS,
                     "traffic": None, "flops_per_launch": flops,
                     "kernel": "picca::cov_syrk_kernel", "kernel_ms": syrk_ms},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        from oracle import cov_oracle as C
        try:
            from threadpoolctl import threadpool_info
            cores = max([p.get("num_threads", 1) for p in threadpool_info()] or [1])
        except ImportError:
            cores = os.cpu_count()
        reps = 0
        tc = time.time()
        while True:
            ref = C.compute_cov(xi, we)
            reps += 1
            if time.time() - tc > 10.0 or reps >= 50:
                break
        cpu_s = (time.time() - tc) / reps
        scale = np.abs(ref).max()
        result["cpu_baseline"] = {
            "value": 1.0 / cpu_s, "unit": "covariances/s", "cores": int(cores),
            "kind": "port", "sample": "the same %d x %d rows, numpy (BLAS) restatement of "
                                      "compute_cov, %d repetitions" % (rows, nb, reps),
            "seconds": cpu_s}
        result["speedup_vs_cpu"] = result["value"] / result["cpu_baseline"]["value"]
        result["parity_sample"] = {
            "max_abs_err_over_max": float(np.abs(cov - ref).max() / scale),
            "symmetric": bool(np.array_equal(cov, cov.T))}
    print(json.dumps(result))


def bench_wick(args):
    """Wick covariance diagrams 1-3 on the 100k-forest survey with the binning of picca's
    own Wick test (15 x 15 bins to 60 Mpc/h): the first --wick-pairs neighbour pairs of
    forests kept by one legacy draw each (reject 0.999), a smooth synthetic weighted 1-D
    correlation per forest (w_k w_l 0.01 exp(-|k - l| / 7), unit variance).  The work per
    pair grows as the square of its in-range pixel pairs, so the pair count is bounded."""
    from picca_amd import batch, synth
    from picca_amd.constants import Cosmo
    from picca_amd.engine import device_info, get_engine, params_from_globals
    from picca_amd.utils import compute_ang_max

    survey = synth.make_survey("survey_100k", seed=args.seed, scale=args.scale)
    store = survey.store
    cosmo = Cosmo(0.315, verbose=False)
    ang_max = compute_ang_max(cosmo, WickGlobals.r_trans_max, survey.z_min)
    pairs = batch.xi_pairs(store, survey.healpix, ang_max)
    keep_forest = np.random.RandomState(args.seed).rand(store.num_forests) > 0.999
    keep = np.flatnonzero(keep_forest[pairs.pair_a])[:args.wick_pairs]
    a, b, ang = pairs.pair_a[keep], pairs.pair_b[keep], pairs.pair_ang[keep]
    log("wick: %d forests kept, the first %d of their neighbour pairs"
        % (int(keep_forest.sum()), len(a)))
    lengths = np.diff(store.offsets)
    wx_off = np.full(store.num_forests, -1, dtype=np.int64)
    mats = []
    used = 0
    for f in np.unique(np.concatenate([a, b])):
        n = lengths[f]
        x = np.arange(n)
        w = store.weights[store.offsets[f]:store.offsets[f + 1]]
        mats.append((np.outer(w, w) * (0.01 * np.exp(-np.abs(x[:, None] - x[None, :]) / 7.0))
                     ).ravel())
        wx_off[f] = used
        used += n * n
    wx = np.concatenate(mats) if mats else np.zeros(0)
    zev = ((1 + store.z) / (1 + WickGlobals.z_ref)) ** (WickGlobals.alpha - 1)
    log("wick: weighted 1-D correlations of %d forests (%d values)" % (len(mats), used))
    params = params_from_globals(WickGlobals)
    eng = get_engine(0)
    info = device_info(0)
    eng.set_forests(store, key="bench_wick")
    if args.traffic_bytes < 0:
        args.traffic_bytes = measured_traffic("wick_100k", args.scale, "wick_kernels")
    if args.copy_probe > 0:
        copy_probe(eng, args.copy_probe)
    for w in range(args.warmup):
        eng.wick_t123(params, a, b, ang, zev, zev, wx_off, wx)
    wall, kernel, rep = [], [], None
    for k in range(args.steps):
        t = time.perf_counter()
        out = eng.wick_t123(params, a, b, ang, zev, zev, wx_off, wx)
        wall.append(time.perf_counter() - t)
        rep = out[5]
        kernel.append(rep["kernel_ms"])
        log("wick step %d: wall %.1f ms, kernels %.1f ms" % (k, wall[-1] * 1e3, kernel[-1]))
    kernel_ms = float(np.mean(kernel))
    products = rep["products"]
    # per (a, b) product of the reference's loop: the category tests aside, two
    # multiplications (T2) or one (T3) and the two symmetric additions
    flops = 3.0 * products
    achieved = flops / (kernel_ms / 1e3) / 1e12
    result = {
        "metric": "Wick pixel-pair products/s (1 GPU)",
        "value": products / (kernel_ms / 1e3), "unit": "pixel-pair products/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * float(np.mean(wall)),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "wick_100k", "scale": args.scale, "forests": store.num_forests,
                   "kept_forests": int(keep_forest.sum()), "forest_pairs": int(len(a)),
                   "reject": 0.999, "bins": "15x15", "r_max": 60.0,
                   "in_range_pixel_pairs": int(rep["in_range_pairs"]),
                   "products": int(products), "batches": int(rep["batches"]),
                   "device": info["name"],
                   "timing": "value: (a, b) products over the kernels' HIP-event time; "
                             "ms_per_step includes the upload of the 1-D correlations"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_VALU_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_VALU_PEAK_TFLOPS,
                     "traffic": args.traffic_bytes if args.traffic_bytes > 0 else None,
                     "flops_per_launch": flops,
                     "kernel": "picca::wick_row_kernel", "kernel_ms": kernel_ms},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        from oracle import cf_oracle as O
        # the first pairs of the workload with pixel pairs in range
        r_par_ok = np.array([
            np.any(np.abs(store.r_comov[store.offsets[x]:store.offsets[x + 1], None]
                          - store.r_comov[None, store.offsets[y]:store.offsets[y + 1]])
                   * np.cos(t / 2) < WickGlobals.r_par_max) for x, y, t in zip(a, b, ang)])
        sel = np.flatnonzero(r_par_ok)[:args.wick_cpu_pairs]
        oparams = O.params_from_module(WickGlobals)
        nb = WickGlobals.num_bins_r_par * WickGlobals.num_bins_r_trans
        ww = np.zeros(nb)
        nw = np.zeros(nb, dtype=np.int64)
        t_ref = [np.zeros(nb * nb) for _ in range(3)]
        sa, sb, sang = a[sel].astype(np.int32), b[sel].astype(np.int32), ang[sel].copy()
        tc = time.time()
        O.lib().oracle_wick_t123(O.ctypes.byref(oparams), len(sel), O._p(sa), O._p(sb),
                                 O._p(sang), O._p(store.offsets), O._p(store.r_comov),
                                 O._p(store.weights), O._p(zev), O._p(zev), O._p(wx_off),
                                 O._p(wx), O._p(ww), O._p(nw), *[O._p(x) for x in t_ref])
        cpu_s = time.time() - tc
        mine = eng.wick_t123(params, sa, sb, sang, zev, zev, wx_off, wx)
        n_ref = float(mine[5]["products"])
        err = max(float(np.abs(m.ravel() - r).max() / max(np.abs(r).max(), 1e-300))
                  for m, r in zip(mine[2:5], t_ref))
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        result["cpu_baseline"] = {
            "value": n_ref / cpu_s, "unit": "pixel-pair products/s", "cores": threads,
            "kind": "port", "sample": "the first %d forest pairs with pixel pairs in range, "
                                      "%.3e products, FP64 C "
                                      "oracle (compute_wickT123_pairs, OpenMP over pairs)"
                                      % (len(sel), n_ref), "seconds": cpu_s}
        if n_ref > 0:
            result["speedup_vs_cpu"] = result["value"] / result["cpu_baseline"]["value"]
        result["parity_sample"] = {"t123_max_abs_err_over_max": err,
                                   "num_pairs_wick_equal": bool(np.array_equal(mine[1], nw))}
    print(json.dumps(result))


def continuum_sets(n, seed, chunk=20000):
    """Forest inputs, binned network inputs and scales of n synthetic spectra, made in
    chunks to bound host memory."""
    from picca_amd import synth_spectra as S
    parts = []
    for k0 in range(0, n, chunk):
        spectra = S.make_spectra(min(chunk, n - k0), seed=seed * 1000003 + k0)
        inputs = S.forest_inputs(spectra)
        x, scale = S.binned_inputs(spectra)
        parts.append((inputs, x, scale))
    offsets = [np.zeros(1, np.int64)]
    base = 0
    for inputs, _, _ in parts:
        offsets.append(inputs["offsets"][1:] + base)
        base += inputs["offsets"][-1]
    merged = {"offsets": np.concatenate(offsets)}
    for name in ("log_lambda", "flux", "ivar", "mean_cont", "x", "eta", "var_lss", "fudge",
                 "log_lambda_min", "log_lambda_max", "z_qso"):
        merged[name] = np.concatenate([p[0][name] for p in parts])
    return merged, np.concatenate([p[1] for p in parts]), np.concatenate([p[2] for p in parts])


def bench_continuum(args):
    """Config 4: continuum fitting of synthetic quasar spectra with picca's DR16 iteration.

    Both sets run the iteration of Dr16ExpectedFlux.compute_expected_flux on the GPU (5
    iterations: batched fit, weighted F / C stack, new mean continuum interpolated to every
    pixel; eta 1, var_lss 0.1, fudge 0 held).  The seed network is trained on the last
    iteration's (a, b) of the training set.  On the test set, with its own final mean
    continuum resident, the last iteration's fit runs from Migrad's start (plain) and from
    the network's prediction (seeded), each under two stopping rules: EDM < 1e-5 (Migrad's
    kind of rule) and EDM < 1e-4 (the final chi2 within about 1e-4 of the minimum); every
    final chi2 is compared with the converged one (EDM < 1e-12 chi2)."""
    from picca_amd import continuum, seed_mlp
    from picca_amd.engine import device_info, get_engine

    n = args.spectra
    iterations = 5
    grid = continuum.rest_frame_grid_log(1040.0, 1200.0, 1e-4)
    t0 = time.time()
    train_in, train_x, train_scale = continuum_sets(n // 2, seed=args.seed + 1)
    test_in, test_x, test_scale = continuum_sets(n, seed=args.seed)
    log("continuum: %d training and %d test spectra, %d + %d forest pixels (%.1f s)"
        % (n // 2, n, train_in["offsets"][-1], test_in["offsets"][-1], time.time() - t0))
    info = device_info(0)
    train = continuum.compute_expected_flux(train_in, grid, iterations, order=1)
    ok = train["fit"]["status"] == 0
    net = seed_mlp.SeedNetwork(seed=args.seed)
    t1 = time.time()
    losses = net.train(train_x[ok], seed_mlp.targets(train["fit"]["a"][ok],
                                                    train["fit"]["b"][ok], train_scale[ok]),
                       epochs=30, batch=512, lr=1e-3)
    train_s = time.time() - t1
    log("seed network: %d Adam steps in %.1f s, loss %.3e -> %.3e"
        % (len(losses), train_s, losses[:20].mean(), losses[-20:].mean()))
    # the test set's own DR16 iteration; its last fit, converged, is the reference minimum
    t2 = time.time()
    test = continuum.compute_expected_flux(test_in, grid, iterations, order=1)
    dr16_wall_s = time.time() - t2
    exact = test["fit"]
    log("DR16 iteration of the test set: %d iterations, %.1f ms of kernels, %.1f s wall"
        % (iterations, test["kernel_ms"], dr16_wall_s))
    eng = get_engine(0)     # holds the test forests and their final mean continuum
    a0, b0 = continuum.start_values(test_in["offsets"], test_in["flux"], test_in["ivar"])
    rules = {}
    for rule, tol_abs in (("edm_1e-5", 1e-5), ("edm_1e-4", 1e-4)):
        plain_ms, seeded_ms, infer_ms = [], [], []
        plain = seeded = None
        for k in range(args.warmup + args.steps):
            plain = eng.cont_fit(a0, b0, tol=1e-15, tol_abs=tol_abs)
            y, inf_ms = net.predict(test_x, timing=True)
            sa, sb = seed_mlp.seeds(y, test_scale)
            seeded = eng.cont_fit(sa, sb, tol=1e-15, tol_abs=tol_abs)
            if k >= args.warmup:
                plain_ms.append(plain["kernel_ms"])
                seeded_ms.append(seeded["kernel_ms"])
                infer_ms.append(inf_ms)
        both = (plain["status"] == 0) & (seeded["status"] == 0) & (exact["status"] == 0)
        d_seeded = np.abs(seeded["chi2"][both] - exact["chi2"][both])
        d_plain = np.abs(plain["chi2"][both] - exact["chi2"][both])
        p_ms = float(np.mean(plain_ms))
        s_ms = float(np.mean(seeded_ms)) + float(np.mean(infer_ms))
        rules[rule] = {
            "plain_fit_ms": p_ms, "seeded_fit_plus_inference_ms": s_ms,
            "inference_kernel_ms": float(np.mean(infer_ms)),
            "seeded_speedup": p_ms / s_ms,
            "plain_fits_per_s": n / (p_ms / 1e3), "seeded_fits_per_s": n / (s_ms / 1e3),
            "plain_mean_passes": float(plain["evaluations"].mean()),
            "seeded_mean_passes": float(seeded["evaluations"].mean()),
            "fits_valid_both": int(both.sum()),
            "chi2_plain_minus_min_max": float(d_plain.max()) if d_plain.size else 0.0,
            "chi2_seeded_minus_min_max": float(d_seeded.max()) if d_seeded.size else 0.0,
            "frac_plain_within_1e-4": float((d_plain < 1e-4).mean()) if d_plain.size else 1.0,
            "frac_seeded_within_1e-4": float((d_seeded < 1e-4).mean()) if d_seeded.size else 1.0,
        }
        log("%s: plain %.3f ms (%.2f passes), seeded %.3f ms (%.2f passes) -> %.2fx"
            % (rule, p_ms, rules[rule]["plain_mean_passes"], s_ms,
               rules[rule]["seeded_mean_passes"], p_ms / s_ms))
    main_rule = rules["edm_1e-4"]
    result = {
        "metric": "continuum fits/s (1 GPU), learned seed vs plain fit",
        "value": main_rule["seeded_fits_per_s"], "unit": "fits/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": main_rule["seeded_fit_plus_inference_ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 fit, bf16/fp32 network",
        "data": "synthetic",
        "config": {"workload": "continuum_%d" % n, "spectra": n, "training_spectra": n // 2,
                   "forest_pixels": int(test_in["offsets"][-1]),
                   "dr16_iterations": iterations,
                   "dr16_iteration_kernel_ms": test["kernel_ms"],
                   "dr16_iteration_wall_s": dr16_wall_s,
                   "rest_frame_bins_kept": int(len(test["x"])),
                   "value_rule": "EDM < 1e-4 (final chi2 within ~1e-4 of the minimum)",
                   "rules": rules,
                   "exact_fit_ms": exact["kernel_ms"],
                   "exact_mean_passes": float(exact["evaluations"].mean()),
                   "training_steps": int(len(losses)), "training_s": train_s,
                   "device": info["name"]},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        from oracle import cont_oracle as C
        from picca_amd import synth_spectra as S
        m = min(n, 300)
        get_mean_cont = C.mean_cont_interpolator(test["x"], test["y"])
        specs = []
        for k in range(m):
            spec = S.forest_spec(test_in, k)
            spec["mean_cont"] = get_mean_cont(spec["log_lambda"] - np.log10(1 + spec["z_qso"]))
            specs.append(spec)
        tc = time.time()
        res = [C.fit(spec, order=1) for spec in specs]
        cpu_s = time.time() - tc
        ref = np.array([r[2] for r in res])
        rel = np.abs(exact["chi2"][:m] - ref) / np.abs(ref)
        result["cpu_baseline"] = {
            "value": m / cpu_s, "unit": "fits/s", "cores": 1, "kind": "port",
            "sample": "first %d test spectra with the final mean continuum, FP64 numpy "
                      "objective minimised by scipy Nelder-Mead + Newton polish (iminuit "
                      "is not installed)" % m,
            "seconds": cpu_s}
        result["parity_sample"] = {"chi2_max_rel_err_vs_oracle": float(rel.max())}
    print(json.dumps(result))


def bench_precision(args):
    """xi of the heavy-tailed survey (delta = 0.3 t(3), w log-uniform in [1e-3, 1e2]) on the
    GPU against the FP64 oracle on a sample of its forest pairs: per-bin relative error and
    edge pairs."""
    args.workload = "survey_100k_heavy"
    args.no_cpu = True
    from picca_amd import batch, synth
    from picca_amd.constants import Cosmo
    from picca_amd.engine import get_engine, params_from_globals
    from picca_amd.utils import compute_ang_max
    from oracle import cf_oracle as O

    survey = synth.make_survey(args.workload, seed=args.seed, scale=args.scale)
    store = survey.store
    ang_max = compute_ang_max(Cosmo(0.315, verbose=False), Globals.r_trans_max, survey.z_min)
    pairs = batch.xi_pairs(store, survey.healpix, ang_max)
    params = params_from_globals(Globals)
    eng = get_engine(0)
    eng.set_forests(store, key="bench_precision")
    eng.set_xi_pairs(pairs.pair_a, pairs.pair_b, pairs.pair_ang, pairs.pair_flags,
                     pairs.pair_row, pairs.num_rows, key="bench_precision")
    full = eng.xi(params)
    sample = sample_pairs(pairs, store, args.cpu_pairs)
    out, stats = O.xi_rows(O.params_from_module(Globals), store, sample.pair_a, sample.pair_b,
                           sample.pair_ang, sample.pair_flags, sample.pair_row,
                           sample.num_rows, num_threads=cpu_cores())
    eng.set_xi_pairs(sample.pair_a, sample.pair_b, sample.pair_ang, sample.pair_flags,
                     sample.pair_row, sample.num_rows, key="sample")
    g = eng.xi(params)
    mask = out[0] > 0
    xi_g = g.xi[mask] / g.weights[mask]
    xi_o = out[1][mask] / out[0][mask]
    rel = np.abs(xi_g - xi_o) / np.abs(xi_o)
    wrel = np.abs(g.weights[mask] - out[0][mask]) / out[0][mask]
    result = {
        "metric": "xi relative error vs FP64 oracle (heavy-tailed delta, w in [1e-3, 1e2])",
        "value": float(rel.max()), "unit": "relative error", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": full.report["kernel_ms"], "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": args.workload, "scale": args.scale,
                   "forest_pairs": pairs.num_pairs, "sample_forest_pairs": sample.num_pairs,
                   "sample_bins": int(mask.sum()),
                   "xi_rel_err_median": float(np.median(rel)),
                   "xi_rel_err_p99": float(np.quantile(rel, 0.99)),
                   "weights_rel_err_max": float(wrel.max()),
                   "num_pairs_mismatches": int(np.count_nonzero(g.num_pairs != out[5])),
                   "edge_pairs_full": int(full.report["edge_pairs"]),
                   "edge_pairs_sample_gpu": int(g.report["edge_pairs"]),
                   "edge_pairs_sample_oracle": int(stats[1]),
                   "overflow_pairs_full": int(full.report["overflow_pairs"])},
    }
    print(json.dumps(result))


class XcfGlobals:
    """picca_xcf / picca_xdmat defaults (py/picca/bin/picca_xcf.py, picca_xdmat.py):
    r_par in [-200, 200) in 100 bins, r_trans < 200 in 50 bins, model binning = data
    binning, quasar redshift-evolution exponent 1.44."""
    num_bins_r_par = 100
    num_bins_r_trans = 50
    num_model_bins_r_par = 100
    num_model_bins_r_trans = 50
    r_par_min = -200.0
    r_par_max = 200.0
    r_trans_max = 200.0
    z_min_pairs = None
    z_max_pairs = None
    zerr_cut_deg = None
    zerr_cut_kms = None
    z_ref = 2.25
    alpha = 2.9
    alpha2 = 1.44
    x_correlation = True
    rmu_binning = False
    ang_correlation = False
    remove_same_half_plate_close_pairs = False
    redshift_evolution_in_distortion_matrix = True


def xcf_workload(args):
    """The 100k-forest survey and its own quasars as the object catalogue (each forest's
    background quasar at z_qso, weight ((1+z)/(1+z_ref))^(1.44-1)), and the forest x
    quasar pairs of fill_neighs (py/picca/xcf.py:71-128): angle < ang_max, not the same
    quasar, r_comov range cut."""
    import types
    from picca_amd import synth
    from picca_amd.constants import Cosmo
    from picca_amd.neighbours import find_neighbours
    from picca_amd.utils import compute_ang_max

    survey = synth.make_survey("survey_100k", seed=args.seed, scale=args.scale)
    store = survey.store
    cosmo = Cosmo(0.315, verbose=False)
    g = XcfGlobals
    qso = types.SimpleNamespace(
        x_cart=store.x_cart, y_cart=store.y_cart, z_cart=store.z_cart, ra=store.ra,
        dec=store.dec, cos_dec=store.cos_dec, thingid=store.thingid, z_qso=store.z_qso)
    qso.r_comov = np.asarray(cosmo.get_r_comov(store.z_qso), dtype=np.float64)
    qso.dist_m = np.asarray(cosmo.get_dist_m(store.z_qso), dtype=np.float64)
    qso.weights = ((1 + store.z_qso) / (1 + g.z_ref)) ** (g.alpha2 - 1)
    z_min2 = float(store.z_qso.min())
    ang_max = compute_ang_max(cosmo, g.r_trans_max, survey.z_min, z_min2)
    order = np.argsort(survey.healpix, kind="stable")
    uniq, counts = np.unique(survey.healpix[order], return_counts=True)
    row_of_src = np.repeat(np.arange(len(uniq), dtype=np.int32), counts)
    t = time.time()
    nl = find_neighbours(store, order, order, ang_max, auto=False, target_store=qso)
    s_of = nl.source_index()
    forest = order[s_of]
    quasar = order[nl.target]
    ang = nl.angle
    first = store.r_comov[store.offsets[:-1]]
    last = store.r_comov[store.offsets[1:] - 1]
    c = np.cos(ang / 2.0)
    keep = ((first[forest] - qso.r_comov[quasar]) * c < g.r_par_max) & \
           ((last[forest] - qso.r_comov[quasar]) * c > g.r_par_min)
    log("xcf neighbours: %d forest x quasar pairs (%.1f s)" % (int(keep.sum()), time.time() - t))
    return (survey, store, qso, forest[keep], quasar[keep], ang[keep], row_of_src[s_of][keep],
            len(uniq), ang_max)


def bench_xcf(args):
    from picca_amd.engine import device_info, get_engine, params_from_globals
    from picca_amd import _native
    from picca_amd.xcf import _quasar_forests

    survey, store, qso, forest, quasar, ang, row, num_rows, ang_max = xcf_workload(args)
    nz = store.nonzero_weight_counts()
    eng = get_engine(0)
    info = device_info(0)
    params = params_from_globals(XcfGlobals)
    if args.mode == "xcf":
        evaluated = int(nz[forest].astype(np.int64).sum())
        eng.set_forests(store, key="bench_xcf")
        eng.xcf_set_pairs(forest, qso.r_comov[quasar], qso.dist_m[quasar], qso.z_qso[quasar],
                          qso.weights[quasar], ang, row, num_rows)
        if args.traffic_bytes < 0:
            args.traffic_bytes = measured_traffic("xcf_100k", args.scale, "xcf_kernel")
        if args.copy_probe > 0:
            copy_probe(eng, args.copy_probe)
        for w in range(args.warmup):
            eng.xcf(params)
        wall, kernel, res = [], [], None
        for k in range(args.steps):
            ts = time.perf_counter()
            res = eng.xcf(params)
            wall.append(time.perf_counter() - ts)
            kernel.append(res.report["kernel_ms"])
            log("xcf step %d: wall %.1f ms, kernel %.1f ms" % (k, wall[-1] * 1e3, kernel[-1]))
        kernel_ms = float(np.mean(kernel))
        rep = res.report
        flops = FLOPS_EVALUATED * evaluated + FLOPS_ACCEPTED_EXTRA * rep["accepted_pairs"]
        achieved = flops / (kernel_ms / 1e3) / 1e12
        result = {
            "metric": "forest x quasar pixel pairs/s (1 GPU)",
            "value": evaluated / (kernel_ms / 1e3), "unit": "pixel pairs/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": 1e3 * float(np.mean(wall)), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "xcf_100k", "scale": args.scale,
                       "forests": store.num_forests, "quasars": store.num_forests,
                       "forest_quasar_pairs": int(len(forest)), "healpix_rows": num_rows,
                       "evaluated_pixel_pairs": evaluated,
                       "accepted_pixel_pairs": int(rep["accepted_pairs"]),
                       "kernel_visited_pixel_pairs": int(rep["visited_pairs"]),
                       "bins": "100x50", "r_par": "[-200, 200)", "r_trans_max": 200.0,
                       "ang_max_deg": float(np.degrees(ang_max)), "device": info["name"],
                       "timing": "value: evaluated pairs over the kernel's HIP-event time; "
                                 "ms_per_step includes the download of the rows"},
            "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_VALU_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": achieved / FP64_VALU_PEAK_TFLOPS,
                         "traffic": (args.traffic_bytes / rep["launches"]
                                     if args.traffic_bytes > 0 else None),
                         "flops_per_launch": flops / rep["launches"],
                         "launches_per_step": int(rep["launches"]),
                         "launch_avg_ms": kernel_ms / rep["launches"],
                         "traffic_per_step": args.traffic_bytes or None,
                         "kernel": "picca::xcf_kernel", "kernel_ms": kernel_ms},
            "cpu_baseline": None,
        }
        if not args.no_cpu:
            from oracle import cf_oracle as O
            oparams = O.params_from_module(XcfGlobals)
            rng = np.random.default_rng(1)
            forests = np.sort(rng.choice(store.num_forests, min(store.num_forests, 1000),
                                         replace=False))
            by_forest = np.argsort(forest, kind="stable")
            starts = np.searchsorted(forest[by_forest], forests)
            ends = np.searchsorted(forest[by_forest], forests, side="right")
            nb = XcfGlobals.num_bins_r_par * XcfGlobals.num_bins_r_trans
            out = [np.zeros(nb) for _ in range(5)] + [np.zeros(nb, dtype=np.int64)]
            stats = np.zeros(3, dtype=np.int64)
            sample_eval = 0
            tc = time.time()
            for f, s0, s1 in zip(forests, starts, ends):
                if s1 == s0:
                    continue
                idx = by_forest[s0:s1]
                q = quasar[idx]
                o0, o1 = store.offsets[f], store.offsets[f + 1]
                O.xcf_pair(oparams, store.z[o0:o1], store.r_comov[o0:o1], store.dist_m[o0:o1],
                           store.weights[o0:o1], store.delta[o0:o1], qso.z_qso[q],
                           qso.r_comov[q], qso.dist_m[q], qso.weights[q], ang[idx], out, stats)
                sample_eval += int(nz[f]) * len(idx)
            cpu_s = time.time() - tc
            result["cpu_baseline"] = {
                "value": sample_eval / cpu_s, "unit": "pixel pairs/s", "cores": 1,
                "kind": "port",
                "sample": "%d random forests (seed 1) with all their quasar neighbours, %.3e "
                          "pixel pairs, FP64 C oracle (scalar restatement of xcf's "
                          "compute_xi_forest_pairs_fast)" % (len(forests), sample_eval),
                "seconds": cpu_s}
            result["speedup_vs_cpu"] = result["value"] / result["cpu_baseline"]["value"]
            # the GPU on the same sample, as one output row: parity with the FP64 oracle
            sel = np.isin(forest, forests)
            eng.xcf_set_pairs(forest[sel], qso.r_comov[quasar[sel]], qso.dist_m[quasar[sel]],
                              qso.z_qso[quasar[sel]], qso.weights[quasar[sel]], ang[sel],
                              np.zeros(int(sel.sum()), dtype=np.int32), 1)
            gres = eng.xcf(params)
            mask = out[0] > 0
            xi_g = gres.xi[0][mask] / gres.weights[0][mask]
            xi_o = out[1][mask] / out[0][mask]
            rel = np.abs(xi_g - xi_o) / np.maximum(np.abs(xi_o), 1e-300)
            result["parity_sample"] = {
                "num_pairs_mismatches": int(np.count_nonzero(gres.num_pairs[0] != out[5])),
                "accepted_pairs_gpu": int(gres.report["accepted_pairs"]),
                "accepted_pairs_oracle": int(stats[0]),
                "edge_pairs_gpu": int(gres.report["edge_pairs"]),
                "edge_pairs_oracle": int(stats[1]),
                "xi_max_rel_err": float(rel.max()) if rel.size else 0.0}
        print(json.dumps(result))
        return

    # distortion matrix: one seeded legacy-numpy draw per forest x quasar pair
    keep = np.random.RandomState(args.seed).rand(len(forest)) > args.reject
    fa, qb, ak = forest[keep], quasar[keep], ang[keep]
    log("xdmat: %d of %d forest x quasar pairs kept" % (len(fa), len(forest)))
    dstore, nf = _quasar_forests(store, qso)
    eng.set_forests(dstore, key="bench_xdmat")
    flags = np.full(len(fa), _native.PICCA_PAIR_XCF_COUNT, dtype=np.uint8)
    for w in range(args.warmup):
        eng.dmat(params, fa, nf + qb, ak, flags)
    times, out = [], None
    for k in range(args.steps):
        t = time.perf_counter()
        out = eng.dmat(params, fa, nf + qb, ak, flags)
        times.append(time.perf_counter() - t)
        log("xdmat step %d: %.2f s" % (k, times[-1]))
    sec = float(np.mean(times))
    stats = out[6]
    result = {
        "metric": "cross distortion-matrix in-range pixel pairs/s (1 GPU)",
        "value": stats[0] / sec, "unit": "pixel pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * sec, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "xdmat_100k", "scale": args.scale,
                   "forests": store.num_forests, "forest_quasar_pairs": int(len(forest)),
                   "kept_pairs": int(len(fa)), "reject": args.reject,
                   "in_range_pairs": int(stats[0]), "selected_pairs": int(stats[1]),
                   "bins": "100x50", "model_bins": "100x50", "device": info["name"],
                   "timing": "wall time of picca_dmat_run, including the 200 MB download of "
                             "the matrix"},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        from oracle import cf_oracle as O
        oparams = O.params_from_module(XcfGlobals)
        rng = np.random.default_rng(1)
        forests = np.unique(fa)
        forests = np.sort(rng.choice(forests, min(len(forests), 40), replace=False))
        nb = XcfGlobals.num_bins_r_par * XcfGlobals.num_bins_r_trans
        nm = XcfGlobals.num_model_bins_r_par * XcfGlobals.num_model_bins_r_trans
        ref = [np.zeros(nb), np.zeros(nb * nm)] + [np.zeros(nm) for _ in range(4)]
        rstats = np.zeros(4, dtype=np.int64)
        sel = np.isin(fa, forests)
        tc = time.time()
        for f in forests:
            m = fa == f
            q = qb[m]
            o0, o1 = store.offsets[f], store.offsets[f + 1]
            O.xdmat_forest(oparams, store.log_lambda[o0:o1], store.r_comov[o0:o1],
                           store.dist_m[o0:o1], store.z[o0:o1], store.weights[o0:o1],
                           qso.r_comov[q], qso.dist_m[q], qso.z_qso[q], qso.weights[q], ak[m],
                           ref, int(store.order[f]), rstats)
        cpu_s = time.time() - tc
        mine = eng.dmat(params, fa[sel], nf + qb[sel], ak[sel], flags[sel])
        err = float(np.abs(mine[1].ravel() - ref[1]).max() / np.abs(ref[1]).max())
        result["cpu_baseline"] = {
            "value": rstats[0] / cpu_s, "unit": "pixel pairs/s", "cores": 1, "kind": "port",
            "sample": "%d random forests (seed 1) with their kept quasars, %d in-range pixel "
                      "pairs, FP64 C oracle (scalar restatement of xcf's "
                      "compute_dmat_forest_pairs_fast)" % (len(forests), int(rstats[0])),
            "seconds": cpu_s}
        result["parity_sample"] = {"dmat_max_abs_err_over_max": err,
                                   "counts_equal": bool(list(mine[6][:2]) == list(rstats[:2]))}
    print(json.dumps(result))


if __name__ == "__main__":
    main()
