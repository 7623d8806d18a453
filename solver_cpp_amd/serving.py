"""Batched and served RVE increments on one GPU (host side of the job server).

``Server`` wraps the library's job server (``sc_server_*``): jobs are RVEs
whose state stays resident in HBM, requests advance a job by one load
increment, admission is by HBM footprint and worker threads overlap the
increments of different jobs on the device.

Workloads (synthetic, seeded):
  batched  - many independent RVEs with cubic grids of uneven edge, random
             grains and orientations, tension along a random direction
             (the load step's rotation R takes x to it), each through its
             whole load case;
  serving  - a Poisson stream of one-increment requests (strain step 1e-4)
             on RVEs of random size, measured for latency and throughput
             against the same requests run one at a time.

    python -m solver_cpp_amd.serving --mode batched|serving [options]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

from .capi import Library, LoadStep, RequestResult, ServerStats
from .geometry import uniform_orientations, voronoi
from .rve import Load, RveSpec, tension_x

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PRODUCT_SO = os.path.join(ROOT, "solver_cpp_amd", "_build", "libsolver_cpp_amd.so")


class Server:
    """A job server of the HIP library."""

    def __init__(self, lib: Library, hbm_budget_bytes: int, workers: int, batched: bool = False,
                 max_wait_ms: float = 5.0):
        """batched: the residual evaluations of the running jobs are computed
        in joint launches (sc_server_create_batched)."""
        self.lib = lib
        h = C.c_void_p()
        if batched:
            lib.check(lib.sc_server_create_batched(int(hbm_budget_bytes), int(workers), float(max_wait_ms),
                                                   C.byref(h)), "sc_server_create_batched")
        else:
            lib.check(lib.sc_server_create(int(hbm_budget_bytes), int(workers), C.byref(h)), "sc_server_create")
        self.h = h
        self._keep = {}

    def add_job(self, spec: RveSpec, loads, keep_resident: bool = False) -> int:
        steps = (LoadStep * len(loads))(*[l.to_c() for l in loads])
        job = C.c_int64()
        self.lib.check(self.lib.sc_server_add_job(self.h, C.byref(spec.desc), steps, len(loads), int(keep_resident),
                                                  C.byref(job)), "sc_server_add_job")
        return job.value

    def submit(self, job: int) -> int:
        r = C.c_int64()
        self.lib.check(self.lib.sc_server_submit(self.h, job, C.byref(r)), "sc_server_submit")
        return r.value

    def wait(self, request: int, timeout_s: float = -1.0) -> RequestResult:
        out = RequestResult()
        rc = self.lib.sc_server_wait(self.h, request, timeout_s, C.byref(out))
        if rc == 4:
            raise TimeoutError(f"request {request} not done after {timeout_s} s")
        self.lib.check(rc, "sc_server_wait")
        return out

    def remove_job(self, job: int):
        self.lib.check(self.lib.sc_server_remove_job(self.h, job), "sc_server_remove_job")

    def field(self, job: int, name: str, n_cells: int) -> np.ndarray:
        """A field of an idle resident job, as Rve.field (sc_server_get_field)."""
        from .capi import SC_FIELD

        n = {"xi": 12, "gamma": 12, "x_p": 3, "iterations": 2}.get(name, 9) * int(n_cells)
        out = np.zeros(n)
        self.lib.check(self.lib.sc_server_get_field(self.h, job, SC_FIELD[name], out.ctypes.data_as(
            C.POINTER(C.c_double)), n), f"sc_server_get_field {name}")
        return out

    def stats(self) -> ServerStats:
        s = ServerStats()
        self.lib.sc_server_get_stats(self.h, C.byref(s))
        return s

    def close(self):
        if self.h:
            self.lib.sc_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rotation_to(u) -> tuple:
    """Unit quaternion (w, x, y, z) of the rotation taking the x axis to u."""
    u = np.asarray(u, dtype=float)
    u = u / np.linalg.norm(u)
    x = np.array([1.0, 0.0, 0.0])
    c = float(x @ u)
    if c < -1 + 1e-12:
        return (0.0, 0.0, 0.0, 1.0)
    axis = np.cross(x, u)
    w = np.sqrt((1.0 + c) / 2.0)
    v = axis / (2.0 * w)
    q = np.array([w, *v])
    return tuple(q / np.linalg.norm(q))


def smooth_edge(e: float) -> int:
    """Nearest integer of the form 2^a 3^b (FFT-friendly grid edge), >= 2."""
    cands = sorted({2 ** a * 3 ** b for a in range(0, 9) for b in range(0, 6) if 2 <= 2 ** a * 3 ** b <= 512})
    return min(cands, key=lambda c: (abs(c - e), c))


def batched_specs(lib, n_rve: int, seed: int = 7, edges=(16, 24, 32, 48), grains=(10, 200), precision: int = 0,
                  **num):
    """Independent RVEs: edge uniform from `edges`, grains uniform in `grains`,
    uniform orientations, tension along a uniform random direction."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_rve):
        e = int(rng.choice(edges))
        g = int(rng.integers(grains[0], grains[1] + 1))
        sub = int(rng.integers(1 << 30))
        cells = (e, e, e)
        size = [1e-5, 1e-5, 1e-5]
        mat = voronoi(cells, size, g, sub)
        ori = uniform_orientations(g, sub + 1)
        u = rng.normal(size=3)
        spec = RveSpec(lib, cells, size, mat, ori, precision=precision, **num)
        out.append((spec, rotation_to(u)))
    return out


def run_batched(lib, specs, workers: int, budget: int, t: float = 5.0, N: int = 10, increments: int = 0,
                batched: bool = False, max_wait_ms: float = 5.0):
    """Every RVE through its load case (the first `increments` of its N
    increments when given) on the server; returns totals.  batched: joint
    launches of the running jobs' residual evaluations."""
    srv = Server(lib, budget, workers, batched=batched, max_wait_ms=max_wait_ms)
    t0 = time.perf_counter()
    reqs = []
    n_inc = increments if increments > 0 else N
    for spec, q in specs:
        job = srv.add_job(spec, [tension_x(t=t, N=N, rot_q=q)], keep_resident=True)
        reqs.append([srv.submit(job) for _ in range(n_inc)])
        srv.remove_job(job)  # memory is released once its last increment is done
    results = [[srv.wait(r) for r in rs] for rs in reqs]
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    updates = sum(r.material_point_updates for rs in results for r in rs)
    bad = sum(1 for rs in results for r in rs if r.status != 0)
    lat = np.array([r.t_end - r.t_start for rs in results for r in rs])
    return dict(wall=wall, updates=updates, value=updates / wall, failed=bad, max_concurrent=st.max_concurrent,
                peak_bytes=st.peak_bytes, results=results, batches=st.batches, max_batch=st.max_batch,
                batched_evaluations=st.batched_evaluations, batch_seconds=st.batch_seconds,
                increment_ms=lat * 1e3)


def serving_requests(lib, n_req: int, rate: float, seed: int = 11, edge_range=(16, 128), precision: int = 0,
                     max_grains: int = 1000, **num):
    """Poisson arrivals at `rate` per second; each request is one increment of
    1e-4 strain (0.1 s at 1e-3/s) on its own RVE."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_req))
    reqs = []
    for i in range(n_req):
        e = smooth_edge(np.exp(rng.uniform(np.log(edge_range[0]), np.log(edge_range[1]))))
        n = e ** 3
        g = int(rng.integers(10, min(max_grains, n // 8) + 1))
        sub = int(rng.integers(1 << 30))
        cells = (e, e, e)
        size = [1e-5, 1e-5, 1e-5]
        spec = RveSpec(lib, cells, size, voronoi(cells, size, g, sub), uniform_orientations(g, sub + 1),
                       precision=precision, **num)
        reqs.append((float(arrivals[i]), spec))
    return reqs


def run_serving(lib, reqs, workers: int, budget: int, batched: bool = False, max_wait_ms: float = 2.0):
    """Submit each request at its arrival time; latency = end - arrival."""
    srv = Server(lib, budget, workers, batched=batched, max_wait_ms=max_wait_ms)
    load = [tension_x(t=0.1, N=1)]
    t0 = time.perf_counter()
    ids = []
    for arr, spec in reqs:
        delay = arr - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        job = srv.add_job(spec, load)
        ids.append((arr, srv.submit(job)))
    res = []
    for arr, r in ids:
        out = srv.wait(r)
        res.append((arr, out))
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    lat = np.array([o.t_end - arr for arr, o in res])  # server clock starts with the server, ~t0
    upd = sum(o.material_point_updates for _, o in res)
    return dict(wall=wall, lat=lat, updates=upd, value=upd / wall, max_concurrent=st.max_concurrent,
                peak_bytes=st.peak_bytes, failed=sum(1 for _, o in res if o.status != 0),
                P_av=[np.array(o.P_av) for _, o in res], batches=st.batches, max_batch=st.max_batch)


def run_one_at_a_time(lib, reqs):
    """The plain loop: each request alone, back to back, one worker."""
    srv = Server(lib, 1 << 62, 1)
    load = [tension_x(t=0.1, N=1)]
    t0 = time.perf_counter()
    lat, P = [], []
    for _, spec in reqs:
        s = time.perf_counter()
        o = srv.wait(srv.submit(srv.add_job(spec, load)))
        lat.append(time.perf_counter() - s)
        P.append(np.array(o.P_av))
    wall = time.perf_counter() - t0
    srv.close()
    return dict(wall=wall, lat=np.array(lat), P_av=P)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--mode", choices=["batched", "serving"], required=True)
    ap.add_argument("--rves", type=int, default=256)
    ap.add_argument("--requests", type=int, default=64)
    ap.add_argument("--rate", type=float, default=32.0, help="requests per second (serving)")
    ap.add_argument("--rates", default="", help="serving: comma-separated rates to sweep (one JSON line each)")
    ap.add_argument("--batched", action="store_true", help="joint launches of the running jobs' evaluations")
    ap.add_argument("--max-wait-ms", type=float, default=2.0, help="batched: longest wait of a request for the others")
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--max-edge", type=int, default=128)
    ap.add_argument("--edges", default="16,24,32,48", help="batched: grid edges drawn from (comma separated)")
    ap.add_argument("--grains", default="10,200", help="batched: grain count range lo,hi")
    ap.add_argument("--load", default="5,10", help="batched: tension duration t and increments N of the load case")
    ap.add_argument("--increments", type=int, default=0, help="batched: increments run per RVE (0: all N)")
    args = ap.parse_args()
    lib = Library(PRODUCT_SO, boundary=True)
    prec = 1 if args.precision == "mixed" else 0
    budget = int(args.budget_gb * (1 << 30))
    num = dict(itmin=4, itmax=40)
    if args.mode == "batched":
        edges = tuple(int(e) for e in args.edges.split(","))
        g_lo, g_hi = (int(g) for g in args.grains.split(","))
        lt, ln = args.load.split(",")
        specs = batched_specs(lib, args.rves, edges=edges, grains=(g_lo, g_hi), precision=prec, **num)
        kw = dict(t=float(lt), N=int(ln), increments=args.increments)
        r = run_batched(lib, specs, args.workers, budget, batched=args.batched, max_wait_ms=args.max_wait_ms, **kw)
        print(f"batched: {args.workers} workers done in {r['wall']:.1f} s", file=sys.stderr, flush=True)
        seq = run_batched(lib, specs, 1, budget, **kw)
        print(json.dumps({"mode": "batched", "rves": args.rves, "workers": args.workers, "precision": args.precision,
                          "joint_launches": args.batched,
                          "edges": list(edges), "grains": [g_lo, g_hi], "load": [float(lt), int(ln)],
                          "increments": args.increments or int(ln),
                          "mat_point_updates_per_s": r["value"], "wall_s": r["wall"],
                          "one_at_a_time_updates_per_s": seq["value"], "speedup": r["value"] / seq["value"],
                          "max_concurrent": r["max_concurrent"], "peak_gb": r["peak_bytes"] / 2 ** 30,
                          "failed_increments": r["failed"]}), flush=True)
    else:
        rates = [float(x) for x in args.rates.split(",")] if args.rates else [args.rate]
        seq = None
        for rate in rates:
            reqs = serving_requests(lib, args.requests, rate, edge_range=(16, args.max_edge), precision=prec,
                                    **num)
            r = run_serving(lib, reqs, args.workers, budget, batched=args.batched, max_wait_ms=args.max_wait_ms)
            print(f"serving at {rate} req/s: {args.workers} workers done in {r['wall']:.1f} s", file=sys.stderr,
                  flush=True)
            if seq is None:  # the same requests (the arrival times aside) at every rate
                seq = run_one_at_a_time(lib, reqs)
            serving_line(args, rate, r, seq)


def serving_line(args, rate, r, seq):
    """One JSON line of a serving run against the same requests one at a time."""
    same = max(float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)) for a, b in zip(r["P_av"], seq["P_av"]))
    print(json.dumps({"mode": "serving", "requests": args.requests, "rate": rate, "workers": args.workers,
                      "joint_launches": args.batched, "batches": r["batches"], "max_batch": r["max_batch"],
                      "precision": args.precision,
                      "p50_ms": float(np.percentile(r["lat"], 50) * 1e3),
                      "p99_ms": float(np.percentile(r["lat"], 99) * 1e3),
                      "mat_point_updates_per_s": r["value"],
                      "one_at_a_time": {"p50_ms": float(np.percentile(seq["lat"], 50) * 1e3),
                                        "p99_ms": float(np.percentile(seq["lat"], 99) * 1e3),
                                        "mat_point_updates_per_s": r["updates"] / seq["wall"]},
                      "max_rel_diff_vs_alone": same, "max_concurrent": r["max_concurrent"],
                      "peak_gb": r["peak_bytes"] / 2 ** 30, "failed": r["failed"]}), flush=True)


if __name__ == "__main__":
    sys.exit(main())
