#!/usr/bin/env python3
"""Throughput and latency of the spectral basic solver on one MI355X.

A step is one load increment of the grid solver (forward + N-GMRES solution +
coordinate update) on a 128^3 periodic polycrystal of 1000 Voronoi grains
(phenopowerlaw FCC, uniaxial tension dot_F_xx = 1e-3/s with the lateral
stresses zero, 60 increments to 3 % strain).  All fields stay resident in HBM;
the timed region covers whole increments.

The metric is material-point updates per second: cells x residual
evaluations over the timed increments / wall time, plus the p50 / p99
increment latency.  One JSON line is printed on stdout.

    python bench.py [--steps K] [--warmup W] [--precision fp64|mixed]
"""
from __future__ import annotations

import argparse
import json
import os
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

# Algorithmic work per unit, per kernel class:
#   response: bytes of the state read and written per cell and evaluation
#             (F F0 Fp0 Lp S0 xi0 gamma0 in, Lp S P Fp Fe xi gamma out, 2 counters)
#             and FP64 flops per Newton iteration / state iteration / cell
#   fft:      9 real planes in, 9 half-spectrum planes out (per cell)
#   gamma:    9 complex values in and out per wavevector
RESPONSE_BYTES_PER_CELL = (9 * 5 + 12 * 2 + 9 * 5 + 12 * 2) * 8 + 8
RESPONSE_FLOPS_NEWTON = 2800  # 8x8 traceless solve: Fe, Hooke, 12 slip rates, residual, assembly 1610, LU 470
RESPONSE_FLOPS_STATE = 850
RESPONSE_FLOPS_CELL = 300
FFT_BYTES_PER_CELL = 9 * 8 + 9 * 16 / 2
GAMMA_BYTES_PER_WAVEVECTOR = 2 * 9 * 16


# HBM bytes per launch of a kernel on this workload, measured by rocprofv3
# FETCH_SIZE / WRITE_SIZE passes over this script and scaled by the same
# counters on a copy kernel of known size; recorded in profiles/
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_128cubed_1000grains.md")


def recorded_traffic(kernel: str):
    """Bytes per launch recorded for `kernel` ('<kernel>_bytes_per_launch: N'), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                key, _, val = line.strip().strip("`").partition(":")
                if key.strip() == f"{kernel}_bytes_per_launch":
                    return float(val)
    except OSError:
        return None
    return None


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


def run_gpu(args, loads, mat, ori, size, precision):
    lib = Library(PRODUCT_SO)
    if lib.backend != "hip-gfx950":
        raise RuntimeError(f"unexpected backend {lib.backend}")
    cells = (args.cells,) * 3
    rve = Rve(lib, cells, size, mat, ori, precision=1 if precision == "mixed" else 0, **EXAMPLE_NUMERICS)
    t0 = time.perf_counter()
    rve.init()
    log(f"[bench] {precision}: init {cells} with {args.grains} grains: {time.perf_counter() - t0:.2f} s")
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
               P=rve.field("P"))
    rve.close()
    return out


def roofline(res):
    kt, st = res["kt"], res["stats"]
    dom = max(kt, key=lambda k: kt[k]["ms"])
    d = kt[dom]
    avg_s = d["ms"] / max(d["launches"], 1) * 1e-3
    per_launch_units = d["points"] / max(d["launches"], 1)
    out = {"kernel": dom, "launches": d["launches"], "avg_ms": d["ms"] / max(d["launches"], 1)}
    if dom == "response":
        flops = (st.newton_iterations * RESPONSE_FLOPS_NEWTON + st.state_iterations * RESPONSE_FLOPS_STATE +
                 d["points"] * RESPONSE_FLOPS_CELL) / max(d["launches"], 1)
        ach = flops / avg_s / 1e12
        out.update(bound="mfma", achieved=ach, peak=FP64_PEAK_TFLOPS, unit="TFLOP/s", frac=ach / FP64_PEAK_TFLOPS,
                   note="FP64 vector work (no MFMA); peak is MI355X's FP64 78.6 TFLOP/s",
                   hbm_gbs=RESPONSE_BYTES_PER_CELL * per_launch_units / avg_s / 1e9)
    else:
        per_unit = {"fft_fwd": FFT_BYTES_PER_CELL, "fft_bwd": FFT_BYTES_PER_CELL,
                    "gamma": GAMMA_BYTES_PER_WAVEVECTOR}.get(dom, 16.0)
        ach = per_unit * per_launch_units / avg_s / 1e9
        out.update(bound="hbm", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s", frac=ach / HBM_PEAK_GBS)
    out["traffic"] = recorded_traffic(dom)
    return out


def cpu_baseline(args, loads):
    """The CPU oracle's full solver on a bounded sample of the same workload:
    a 32^3 grid with the same grain size (grains scaled by the volume) through
    the first increments of the same load case."""
    lib = Library(ORACLE_SO)
    c = args.cpu_sample_cells
    grains = max(2, round(args.grains * (c / args.cells) ** 3))
    mat, ori, size = polycrystal((c, c, c), grains, seed=args.seed, ori_seed=args.ori_seed)
    rve = Rve(lib, (c, c, c), size, mat, ori, **EXAMPLE_NUMERICS)
    rve.init()
    it = rve.increments(loads)
    t0 = time.perf_counter()
    evals, n_inc = 0, 0
    while n_inc < args.cpu_sample_increments:
        r = next(it)
        evals += r["evaluations"]
        n_inc += 1
        log(f"[bench] cpu oracle increment {r['increment']}: {r['seconds']:.2f} s, {r['evaluations']} evaluations")
    wall = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": c ** 3 * evals / wall, "unit": "mat-point updates/s", "cores": threads, "kind": "port",
            "sample": f"oracle (FP64 C++ restatement, OpenMP) full solver, {c}^3 grid, {grains} grains, "
                      f"first {n_inc} increments of the same load case ({evals} residual evaluations, {wall:.1f} s)"}


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
    value = res["N"] * res["evals"] / res["wall"]
    plain_value = plain["N"] * plain["evals"] / plain["wall"]
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
                       "residual_evaluations": plain["evals"]},
        "speedup_vs_plain_fp64": value / plain_value,
        "increment_time_ratio_plain_over_this": float(plain_ms.mean() / lat_ms.mean()),
        "error_vs_plain_fp64": {
            "P_av_rel": float(np.abs(res["P_av"] - plain["P_av"]).max() / np.abs(plain["P_av"]).max()),
            "P_max_rel": float(np.abs(res["P"] - plain["P"]).max() / np.abs(plain["P"]).max())},
        "fp32_evaluations": int(res["stats"].fp32_evaluations),
        "residual_evaluations": res["evals"],
        "newton_iterations": int(res["stats"].newton_iterations),
        "state_iterations": int(res["stats"].state_iterations),
        "kernels_ms": {k: round(v["ms"], 3) for k, v in res["kt"].items() if v["launches"]},
        "roofline": roofline(res),
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args, loads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
