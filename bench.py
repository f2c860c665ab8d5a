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
FP64_VALU_PEAK_TFLOPS = 78.6   # MI355X FP64 vector, published
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


def measured_traffic(workload, scale):
    """HBM bytes per pair-kernel launch measured with FETCH_SIZE + WRITE_SIZE (each scaled
    by the same counters on a plain copy), as recorded beside the profiles."""
    path = os.path.join(ROOT, "profiles", "xi_kernel_hbm_traffic.json")
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
    parser.add_argument("--cpu-pairs", type=float, default=1.5e10,
                        help="evaluated pixel pairs in the CPU-baseline sample")
    parser.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    parser.add_argument("--copy-probe", type=float, default=0.0,
                        help="MiB of a plain device copy to run before the steps")
    parser.add_argument("--traffic-bytes", type=float, default=-1.0,
                        help="HBM bytes per launch of the pair kernel (default: the "
                             "counter-derived figure in profiles/ for this workload)")
    args = parser.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: this engine has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

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

    for w in range(args.warmup):
        rep = eng.xi_execute(params)
        log("warmup %d: kernel %.1f ms" % (w, rep.kernel_ms))
    wall = []
    kernel = []
    report = None
    for k in range(args.steps):
        ts = time.perf_counter()
        report = eng.xi_execute(params)   # returns after the stream has drained
        wall.append(time.perf_counter() - ts)
        kernel.append(report.kernel_ms)
        log("step %d: wall %.1f ms, kernel %.1f ms" % (k, wall[-1] * 1e3, report.kernel_ms))
    ms_per_step = 1e3 * float(np.mean(wall))
    kernel_ms = float(np.mean(kernel))
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
                "kernel": "picca::xi_kernel<0,false,false,true>", "kernel_ms": kernel_ms,
                "hbm_gbs_at_traffic": (args.traffic_bytes / (kernel_ms / 1e3) / 1e9
                                       if args.traffic_bytes > 0 else None)}

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "pixel pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
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
            "edge_pairs": rep["edge_pairs"],
            "overflow_pairs": rep["overflow_pairs"],
            "bins": "50x50", "r_par_max": 200.0, "r_trans_max": 200.0, "nside": survey.nside,
            "z_min": survey.z_min, "ang_max_deg": float(np.degrees(ang_max)),
            "neighbour_search_s": neighbour_s, "upload_s": upload_s,
            "device": info["name"],
        },
        "roofline": roofline,
        "cpu_baseline": None,
    }

    if not args.no_cpu:
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
        # GPU on the same sample: parity with the FP64 oracle
        eng.set_xi_pairs(sample.pair_a, sample.pair_b, sample.pair_ang, sample.pair_flags,
                         sample.pair_row, sample.num_rows, key="sample")
        gres = eng.xi(params)
        nb_mismatch = int(np.count_nonzero(gres.num_pairs != out[5]))
        we_o = out[0]
        mask = we_o > 0
        xi_g = gres.xi[mask] / gres.weights[mask]
        xi_o = out[1][mask] / we_o[mask]
        rel = np.abs(xi_g - xi_o) / np.maximum(np.abs(xi_o), 1e-300)
        result["cpu_baseline"] = {
            "value": cpu_rate, "unit": "pixel pairs/s", "cores": cores, "kind": "port",
            "sample": "%d random forest pairs (seed 1) of the same workload, %.3e pixel pairs, "
                      "FP64 C oracle (scalar restatement of the reference loop), one OpenMP "
                      "task per healpix row" % (sample.num_pairs, sample_eval),
            "seconds": cpu_s,
        }
        result["parity_sample"] = {
            "num_pairs_mismatches": nb_mismatch,
            "accepted_pairs_gpu": int(gres.report["accepted_pairs"]),
            "accepted_pairs_oracle": int(stats[0]),
            "edge_pairs_gpu": int(gres.report["edge_pairs"]),
            "edge_pairs_oracle": int(stats[1]),
            "xi_max_rel_err": float(rel.max()) if rel.size else 0.0,
            "xi_median_rel_err": float(np.median(rel)) if rel.size else 0.0,
        }
        result["speedup_vs_cpu"] = value / cpu_rate
    print(json.dumps(result))


if __name__ == "__main__":
    main()
