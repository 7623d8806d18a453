#!/usr/bin/env python3
"""Throughput and latency of the spectral basic solver on one MI355X.

A step is one load increment of the grid solver (forward + N-GMRES solution +
coordinate update) on a 128^3 periodic polycrystal of 1000 Voronoi grains
(phenopowerlaw FCC, uniaxial tension dot_F_xx = 1e-3/s with the lateral
stresses zero, 60 increments to 3 % strain).  All fields stay resident in HBM;
the timed region covers whole increments.

The metric is material-point updates per second: cells x material-point
updates (constitutive evaluations) over the timed increments / wall time,
plus the p50 / p99 increment latency.  The plain FP64 form runs the same
increments for the comparison; the equal-work comparison is the increment
time.  A second part runs the batched workload (256 RVEs of uneven grids
through their load cases) through the batched job server, in both
precisions, against the same RVEs run one at a time in FP64, and a third
the serving stream (Poisson requests of one increment on RVEs of edges 16 to
128) at several arrival rates.  One JSON line is printed on stdout.

    python bench.py [--steps K] [--warmup W] [--precision fp64|mixed]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from solver_cpp_amd.capi import Library  # noqa: E402
from solver_cpp_amd.geometry import CONFIGS, EXAMPLE_NUMERICS, polycrystal  # noqa: E402
from solver_cpp_amd.rve import Rve, tension_x  # noqa: E402

PRODUCT_SO = os.path.join(ROOT, "solver_cpp_amd", "_build", "libsolver_cpp_amd.so")
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

METRIC = "mat-point updates/s @128^3 vs plain FP64 GPU; p50/p99 increment latency"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# Algorithmic work of the material-point update (own count of this
# formulation; the kernel counts its Newton and state iterations): FP64
# flops per Newton iteration (Fe, Hooke, 12 resolved shears and slip rates,
# residual, the 8x8 Jacobian from the Hadamard sums of the slip groups and
# its LU, as the compiled loop body executes them), per state iteration
# (the rest of integrateStress, dotState with its quarter powers, the damped
# update) and per cell (entry, restore, first rates, stores), and bytes of
# state read and written per cell and evaluation (F F0 Fp0 Lp S0 xi0 gamma0
# in, Lp S P Fp Fe xi gamma out, 2 counters).  The FP64 instructions
# rocprofv3 counts over the timed launches of this workload (x 64 lanes x
# the active-lane share) differ from this count (line-search iterations
# skip the Jacobian); the ratio is recorded in profiles/ and scales the count.
RESPONSE_FLOPS_NEWTON = 1500
RESPONSE_FLOPS_STATE = 1000
RESPONSE_FLOPS_CELL = 500
RESPONSE_BYTES_PER_CELL = (9 * 5 + 12 * 2 + 9 * 5 + 12 * 2) * 8 + 8

# HBM bytes per launch of the response kernel on this workload, measured by
# rocprofv3 FETCH_SIZE / WRITE_SIZE passes over this script and scaled by the
# same counters on a copy kernel of known size; recorded in profiles/
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_128cubed_1000grains.md")
COUNTERS_FILE = os.path.join(ROOT, "profiles", "response_kernel_counters_128cubed_1000grains.md")


def recorded(path: str, key: str):
    """The value of a '`key: N`' line of a profile summary, or None."""
    try:
        with open(path) as f:
            for line in f:
                k, _, val = line.strip().strip("`").partition(":")
                if k.strip() == key:
                    return float(val)
    except OSError:
        return None
    return None


def recorded_traffic(kernel: str):
    """HBM bytes per launch recorded for `kernel`, or None."""
    return recorded(TRAFFIC_FILE, f"{kernel}_bytes_per_launch")


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


def run_gpu(args, loads, mat, ori, size, precision, **num):
    lib = Library(PRODUCT_SO)
    if lib.backend != "hip-gfx950":
        raise RuntimeError(f"unexpected backend {lib.backend}")
    cells = (args.cells,) * 3
    numerics = dict(EXAMPLE_NUMERICS)
    numerics.update(num)
    rve = Rve(lib, cells, size, mat, ori, precision=1 if precision == "mixed" else 0, **numerics)
    t0 = time.perf_counter()
    rve.init()
    log(f"[bench] {precision}{' tight' if num else ''}: init {cells} with {args.grains} grains: "
        f"{time.perf_counter() - t0:.2f} s")
    it = rve.increments(loads)
    for _ in range(args.warmup):
        r = next(it)
        log(f"[bench] warmup increment {r['increment']}: {r['seconds'] * 1e3:.1f} ms, "
            f"{r['iterations']} iterations, {r['evaluations']} evaluations")
    rve.reset_stats()
    # HIP events around the material-point step only (the roofline's kernel):
    # timing every class costs the timed region about 2 %
    rve.profile(True, classes=["response"])
    lat, evals = [], 0
    t_start = time.perf_counter()
    for _ in range(args.steps):
        r = next(it)
        lat.append(r["seconds"])
        evals += r["evaluations"]
        log(f"[bench] increment {r['increment']}: {r['seconds'] * 1e3:.1f} ms, "
            f"{r['iterations']} iterations, {r['evaluations']} evaluations")
    wall = time.perf_counter() - t_start
    kt = rve.kernel_times()
    st = rve.stats()
    rve.profile(False)
    N = int(np.prod(cells))
    out = dict(N=N, wall=wall, lat=lat, evals=evals, kt=kt, stats=st, P_av=np.array(rve.scalars().P_av),
               P=rve.field("P"), updates=int(st.material_point_updates))
    rve.close()
    return out


def roofline(res):
    """The response kernel (dominant): algorithmic FP64 flops per launch over
    its average HIP-event time on the solver's stream."""
    kt, st = res["kt"], res["stats"]
    d = kt["response"]
    launches = max(d["launches"], 1)
    avg_s = d["ms"] / launches * 1e-3
    cells = d["points"] / launches
    by_count = (st.newton_iterations * RESPONSE_FLOPS_NEWTON + st.state_iterations * RESPONSE_FLOPS_STATE +
                d["points"] * RESPONSE_FLOPS_CELL) / launches
    calib = recorded(COUNTERS_FILE, "response_flops_counter_over_count") or 1.0
    flops = by_count * calib
    ach = flops / avg_s / 1e12
    return {"kernel": "response", "launches": d["launches"], "avg_ms": d["ms"] / launches,
            "bound": "fp64-valu", "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": ach / FP64_PEAK_TFLOPS,
            "note": "FP64 vector ALU work (no dense contraction, no MFMA): peak is MI355X's FP64 vector "
                    "78.6 TFLOP/s; flops: the kernel's own Newton/state iteration counts x per-iteration "
                    "counts, scaled by the executed-FP64-instruction ratio rocprofv3 recorded for this workload",
            "flops_per_launch": flops, "flops_per_launch_by_count": by_count, "flop_calibration": calib,
            "cells_per_launch": cells,
            "hbm_gbs_algorithmic": RESPONSE_BYTES_PER_CELL * cells / avg_s / 1e9,
            "traffic": recorded_traffic("response")}


def run_batched_workload(args, precision: str, batched: bool, workers: int):
    from solver_cpp_amd.serving import batched_specs, run_batched

    lib = Library(PRODUCT_SO, boundary=True)
    prec = 1 if precision == "mixed" else 0
    specs = batched_specs(lib, args.batched_rves, seed=7, precision=prec, **EXAMPLE_NUMERICS)
    t0 = time.perf_counter()
    r = run_batched(lib, specs, workers, int(200 * (1 << 30)), t=5.0, N=10, batched=batched, max_wait_ms=5.0)
    log(f"[bench] batched workload ({precision}, {'joint launches' if batched else 'one at a time'}): "
        f"{r['value']:.3e} updates/s, {r['wall']:.1f} s, batches {r['batches']}, max batch {r['max_batch']}, "
        f"failed {r['failed']}")
    out = {"value": r["value"], "wall_s": r["wall"], "updates": r["updates"], "failed_increments": r["failed"],
           "p50_increment_ms": float(np.percentile(r["increment_ms"], 50)),
           "p99_increment_ms": float(np.percentile(r["increment_ms"], 99)),
           "peak_gb": r["peak_bytes"] / 2 ** 30}
    if batched:
        out.update(batches=r["batches"], max_batch=r["max_batch"],
                   mean_batch=r["batched_evaluations"] / max(r["batches"], 1),
                   joint_launch_seconds=r["batch_seconds"])
    P = [np.array(x.P_av) for rs in r["results"] for x in rs]
    del specs
    return out, P


def run_serving_sweep(args):
    """The serving stream: Poisson requests, each one 1e-4 strain increment on
    its own RVE (edges log-uniform in [16, 128], rounded to FFT-friendly
    lengths), through the batched job server at several arrival rates, and
    the same requests one at a time (one worker, no joint launches)."""
    from solver_cpp_amd.serving import run_one_at_a_time, run_serving, serving_requests

    lib = Library(PRODUCT_SO, boundary=True)
    rates = [float(x) for x in args.serving_rates.split(",")]
    out = {"workload": f"{args.serving_requests} Poisson requests (generator seed 11), each one increment of "
                       f"1e-4 strain (tension_x, 0.1 s) on its own RVE: cubic grid, edge log-uniform in [16, 128] "
                       f"rounded to 2^a 3^b, 10..min(1000, cells/8) Voronoi grains; batched server, 16 workers",
           "rates": []}
    seq = None
    for rate in rates:
        reqs = serving_requests(lib, args.serving_requests, rate, **EXAMPLE_NUMERICS)
        r = run_serving(lib, reqs, 16, int(200 * (1 << 30)), batched=True, max_wait_ms=2.0)
        if seq is None:
            seq = run_one_at_a_time(lib, reqs)
            out["one_at_a_time"] = {"p50_ms": float(np.percentile(seq["lat"], 50) * 1e3),
                                    "p99_ms": float(np.percentile(seq["lat"], 99) * 1e3),
                                    "mat_point_updates_per_s": r["updates"] / seq["wall"]}
        same = all(np.array_equal(a, b) for a, b in zip(r["P_av"], seq["P_av"]))
        row = {"rate_per_s": rate, "p50_ms": float(np.percentile(r["lat"], 50) * 1e3),
               "p99_ms": float(np.percentile(r["lat"], 99) * 1e3), "mat_point_updates_per_s": r["value"],
               "max_concurrent": r["max_concurrent"], "batches": r["batches"], "max_batch": r["max_batch"],
               "failed": r["failed"], "bitwise_equal_to_one_at_a_time": bool(same)}
        log(f"[bench] serving at {rate:g} req/s: p50 {row['p50_ms']:.1f} ms, p99 {row['p99_ms']:.1f} ms, "
            f"{row['mat_point_updates_per_s']:.3e} updates/s, max concurrent {row['max_concurrent']}")
        out["rates"].append(row)
        del reqs
    return out


def cpu_info():
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        model = platform.processor()
    nproc = os.cpu_count() or 1
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=10).stdout.strip())
    except Exception:
        pass
    return model, nproc


def _omp_threads(n: int):
    """Thread count of the oracle's OpenMP loops (the material point)."""
    try:
        C.CDLL("libgomp.so.1").omp_set_num_threads(int(n))
    except OSError:
        pass


def _oracle_sample(lib, c, grains, args, loads, n_inc):
    mat, ori, size = polycrystal((c, c, c), grains, seed=args.seed, ori_seed=args.ori_seed)
    rve = Rve(lib, (c, c, c), size, mat, ori, **EXAMPLE_NUMERICS)
    rve.init()
    it = rve.increments(loads)
    t0 = time.perf_counter()
    evals = 0
    for _ in range(n_inc):
        r = next(it)
        evals += r["evaluations"]
    wall = time.perf_counter() - t0
    rve.close()
    return c ** 3 * evals / wall, evals, wall


def cpu_baseline(args, loads, timed_evals: int, gpu_wall: float):
    """The CPU oracle (the FP64 C++ restatement, OpenMP over material points)
    on bounded samples of the same workload, with the same grain size (grains
    scaled by the volume) through the first increments of the same load case:
    all threads on a 32^3 sample (the reported value), one thread on a 16^3
    sample, and one residual evaluation of the 128^3 workload itself (all
    threads) from which the full run's CPU time is extrapolated."""
    lib = Library(ORACLE_SO)
    model, nproc = cpu_info()
    threads = int(os.environ.get("OMP_NUM_THREADS", nproc))

    def grains_for(c):
        return max(2, round(args.grains * (c / args.cells) ** 3))

    _omp_threads(threads)
    v_all, e_all, w_all = _oracle_sample(lib, args.cpu_sample_cells, grains_for(args.cpu_sample_cells), args, loads,
                                         args.cpu_sample_increments)
    log(f"[bench] cpu oracle {threads} threads, {args.cpu_sample_cells}^3: {v_all:.3e} updates/s ({w_all:.1f} s)")
    _omp_threads(1)
    v_one, e_one, w_one = _oracle_sample(lib, 16, grains_for(16), args, loads, args.cpu_sample_increments)
    log(f"[bench] cpu oracle 1 thread, 16^3: {v_one:.3e} updates/s ({w_one:.1f} s)")
    _omp_threads(threads)
    out = {"value": v_all, "unit": "mat-point updates/s", "cores": threads, "kind": "port",
           "sample": f"oracle (FP64 C++ restatement; OpenMP over material points, FFTs serial) full solver, "
                     f"{args.cpu_sample_cells}^3 grid, {grains_for(args.cpu_sample_cells)} grains, first "
                     f"{args.cpu_sample_increments} increments of the same load case ({e_all} residual "
                     f"evaluations, {w_all:.1f} s)",
           "one_thread": {"value": v_one, "sample": f"16^3 grid, {grains_for(16)} grains, first "
                                                    f"{args.cpu_sample_increments} increments ({e_one} evaluations, "
                                                    f"{w_one:.1f} s)"},
           "cpu_model": model, "nproc": nproc}
    if not args.no_cpu_128:
        c = args.cells
        mat, ori, size = polycrystal((c, c, c), args.grains, seed=args.seed, ori_seed=args.ori_seed)
        rve = Rve(lib, (c, c, c), size, mat, ori, **EXAMPLE_NUMERICS)
        rve.init()
        rve.forward(False, False, loads[0].t / loads[0].N, 1.0, loads[0].t, loads[0])
        F = rve.field("F")
        t0 = time.perf_counter()
        rve.form_residual(F)
        t_eval = time.perf_counter() - t0
        rve.close()
        out["workload_evaluation"] = {
            "cells": [c, c, c], "seconds": t_eval, "updates_per_s": c ** 3 / t_eval, "threads": threads,
            "timed_increments_extrapolated_s": timed_evals * t_eval, "gpu_plain_fp64_timed_s": gpu_wall,
            "extrapolation": f"the timed increments' {timed_evals} residual evaluations (plain FP64 form, the GPU's "
                             f"count; the oracle takes the same iterations) x this evaluation's time"}
        log(f"[bench] cpu oracle one {c}^3 evaluation: {t_eval:.1f} s")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed increments")
    ap.add_argument("--warmup", type=int, default=2, help="untimed increments before the timed ones")
    c3 = CONFIGS["C3"]
    ap.add_argument("--cells", type=int, default=c3["edge"])
    ap.add_argument("--grains", type=int, default=c3["grains"])
    ap.add_argument("--increments", type=int, default=c3["N"], help="increments of the load case (3 %% strain)")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed",
                    help="measured form; the plain FP64 form runs the same increments for the comparison")
    ap.add_argument("--seed", type=int, default=c3["seed"], help="Voronoi seed")
    ap.add_argument("--ori-seed", type=int, default=c3["ori_seed"], help="orientation seed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-cells", type=int, default=32)
    ap.add_argument("--cpu-sample-increments", type=int, default=2)
    ap.add_argument("--no-cpu-128", action="store_true", help="skip the CPU time of one residual evaluation of the "
                                                                 "workload itself")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched workload")
    ap.add_argument("--batched-rves", type=int, default=256)
    ap.add_argument("--no-tight", action="store_true", help="skip the tight-tolerance precision comparison")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving sweep")
    ap.add_argument("--serving-rates", default="32,128,512", help="serving sweep: arrival rates per second")
    ap.add_argument("--serving-requests", type=int, default=64)
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU: the solver path is built and measured on a single MI355X; "
              "several GPUs are not supported by this benchmark.", file=sys.stderr)
        sys.exit(2)
    if args.warmup + args.steps > args.increments:
        sys.exit("warmup + steps exceeds the increments of the load case")

    rate = 1.0e-3
    t_total = 0.03 / rate  # 3 % strain
    loads = [tension_x(t=t_total, N=args.increments, rate=rate)]
    cells = (args.cells,) * 3
    mat, ori, size = polycrystal(cells, args.grains, seed=args.seed, ori_seed=args.ori_seed)
    plain = run_gpu(args, loads, mat, ori, size, "fp64")
    res = plain if args.precision == "fp64" else run_gpu(args, loads, mat, ori, size, args.precision)
    value = res["updates"] / res["wall"]
    plain_value = plain["updates"] / plain["wall"]
    plain_evals, plain_wall = int(plain["evals"]), float(plain["wall"])
    lat_ms = np.array(res["lat"]) * 1e3
    plain_ms = np.array(plain["lat"]) * 1e3
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "mat-point updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": float(lat_ms.mean()),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"spectral basic grid solver, {args.cells}^3 periodic grid (1e-5 m cube), {args.grains} "
                        f"Voronoi grains (seeds {args.seed}/{args.ori_seed}), phenopowerlaw FCC, uniaxial tension "
                        f"dot_F_xx = 1e-3/s, {args.increments} increments to 3 % strain, N_iter_min/max = 4/40",
            "cells": list(cells),
            "grains": args.grains,
            "precision": args.precision,
            "timed_increments": [args.warmup + 1, args.warmup + args.steps],
        },
        "p50_increment_ms": float(np.percentile(lat_ms, 50)),
        "p99_increment_ms": float(np.percentile(lat_ms, 99)),
        "plain_fp64": {"value": plain_value, "ms_per_step": float(plain_ms.mean()),
                       "p50_increment_ms": float(np.percentile(plain_ms, 50)),
                       "p99_increment_ms": float(np.percentile(plain_ms, 99)),
                       "residual_evaluations": plain["evals"], "material_point_updates": plain["updates"]},
        "speedup_vs_plain_fp64": value / plain_value,
        # the equal-work comparison: both forms solve the same increments
        "increment_time_ratio_plain_over_this": float(plain_ms.mean() / lat_ms.mean()),
        "error_vs_plain_fp64": {
            "P_av_rel": float(np.abs(res["P_av"] - plain["P_av"]).max() / np.abs(plain["P_av"]).max()),
            "P_max_rel": float(np.abs(res["P"] - plain["P"]).max() / np.abs(plain["P"]).max())},
        "fp32_evaluations": int(res["stats"].fp32_evaluations),
        "residual_evaluations": res["evals"],
        "material_point_updates": res["updates"],
        "newton_iterations": int(res["stats"].newton_iterations),
        "state_iterations": int(res["stats"].state_iterations),
        "kernels_ms": {k: round(v["ms"], 3) for k, v in res["kt"].items() if v["launches"]},
        "roofline": roofline(res),
    }
    del plain, res
    if args.precision == "mixed" and not args.no_tight:
        # precision error apart from the stopping spread: both forms with the
        # divergence and stress tolerances 1000 times tighter
        tight = dict(itmax=100, eps_rel_div=5e-7, eps_abs_div=1e-12, eps_rel_P=1e-6, eps_abs_P=1e-6)
        targs = argparse.Namespace(**vars(args))
        targs.steps, targs.warmup = 2, 1
        tp = run_gpu(targs, loads, mat, ori, size, "fp64", **tight)
        tm = run_gpu(targs, loads, mat, ori, size, "mixed", **tight)
        line["error_vs_plain_fp64_tight_tolerances"] = {
            "P_av_rel": float(np.abs(tm["P_av"] - tp["P_av"]).max() / np.abs(tp["P_av"]).max()),
            "P_max_rel": float(np.abs(tm["P"] - tp["P"]).max() / np.abs(tp["P"]).max()),
            "tolerances": tight, "increments": [2, 3]}
        del tp, tm
    if not args.no_batched:
        bm, Pm = run_batched_workload(args, "mixed", True, args.batched_rves)
        bf, Pf = run_batched_workload(args, "fp64", True, args.batched_rves)
        one, Po = run_batched_workload(args, "fp64", False, 1)
        same = all(np.array_equal(a, b) for a, b in zip(Pf, Po))
        err = max(float(np.abs(a - b).max() / np.abs(b).max()) for a, b in zip(Pm, Po))
        line["batched"] = {
            "workload": f"{args.batched_rves} RVEs (generator seed 7): cubic grids with edges uniform in "
                        f"{{16,24,32,48}}, 10-200 Voronoi grains, tension along uniform random directions, 10 "
                        f"increments of 0.5 s each; one material-point launch per evaluation for all running RVEs",
            "mixed": bm, "fp64": bf, "plain_fp64_one_at_a_time": one,
            "speedup_mixed_vs_plain_one_at_a_time": bm["value"] / one["value"],
            "speedup_fp64_vs_plain_one_at_a_time": bf["value"] / one["value"],
            "fp64_bitwise_equal_to_one_at_a_time": bool(same),
            "mixed_max_P_av_rel_diff_vs_fp64": err}
    if not args.no_serving:
        line["serving"] = run_serving_sweep(args)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args, loads, plain_evals, plain_wall)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
