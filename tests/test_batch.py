"""Batched residual evaluations (sc_server_create_batched): the evaluations
of all running jobs are computed in joint launches -- one material-point
launch over every job's cells, one batched FFT per pass and one Gamma launch
per group of equal grids.  Each job's arithmetic is unchanged, so every
result must be bitwise the one of the same job run alone on its own stream,
and the workload's oracle fixtures hold for the batched form as well."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import GOLDEN, PRODUCT_SO, tail_close
from solver_cpp_amd.capi import Library
from solver_cpp_amd.geometry import EXAMPLE_NUMERICS
from solver_cpp_amd.rve import tension_x
from solver_cpp_amd.serving import Server, batched_specs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    return Library(PRODUCT_SO, boundary=True)


def _run(lib, specs, loads, incs, batched, workers):
    srv = Server(lib, 1 << 37, workers, batched=batched, max_wait_ms=20.0)
    jobs = [srv.add_job(s, l, keep_resident=True) for (s, _), l in zip(specs, loads)]
    reqs = [[srv.submit(j) for _ in range(incs)] for j in jobs]
    res = [[srv.wait(r, 300.0) for r in rs] for rs in reqs]
    fields = [{k: srv.field(j, k, s.N) for k in ("F", "P", "Fp", "xi")} for j, (s, _) in zip(jobs, specs)]
    st = srv.stats()
    for j in jobs:
        srv.remove_job(j)
    srv.close()
    return res, fields, st


@pytest.mark.parametrize("precision", [0, 1])
def test_batched_equals_alone(lib, precision):
    """Eight RVEs of uneven grids (8, 12, 16, two of them of one size) and
    random tension directions, three increments each: alone (one worker, one
    job at a time) and batched (eight workers) give bitwise the same counts,
    <P> and fields, and the batched run did join evaluations."""
    specs = batched_specs(lib, 8, seed=9, edges=(8, 12, 16), grains=(3, 12), precision=precision,
                          **EXAMPLE_NUMERICS)
    loads = [[tension_x(t=1.5, N=3, rot_q=q)] for _, q in specs]
    print(f"[batch] precision {precision}: edges {[s.cells[0] for s, _ in specs]}", flush=True)
    alone, fa, _ = _run(lib, specs, loads, 3, batched=False, workers=1)
    batch, fb, st = _run(lib, specs, loads, 3, batched=True, workers=8)
    assert st.batches > 0 and st.max_batch >= 2, (st.batches, st.max_batch)
    for a_job, b_job in zip(alone, batch):
        for a, b in zip(a_job, b_job):
            assert a.status == 0 and b.status == 0
            assert (a.iterations, a.evaluations, a.increment) == (b.iterations, b.evaluations, b.increment)
            np.testing.assert_array_equal(np.array(a.P_av), np.array(b.P_av))
    for x, y in zip(fa, fb):
        for k in x:
            np.testing.assert_array_equal(x[k], y[k], err_msg=k)


def test_batched_odd_grid_equals_alone(lib):
    """Grids whose x pass pairs rows across an odd row count (27^3: 9 x 27 x
    27 rows per RVE) joined in one transform: each RVE's rows are paired
    among themselves, so the batched results are bitwise those alone."""
    specs = batched_specs(lib, 4, seed=13, edges=(27,), grains=(4, 8), **EXAMPLE_NUMERICS)
    loads = [[tension_x(t=1.0, N=2, rot_q=q)] for _, q in specs]
    alone, fa, _ = _run(lib, specs, loads, 2, batched=False, workers=1)
    batch, fb, st = _run(lib, specs, loads, 2, batched=True, workers=4)
    assert st.max_batch >= 2, st.max_batch
    for a_job, b_job in zip(alone, batch):
        for a, b in zip(a_job, b_job):
            assert a.status == 0 and b.status == 0
            assert (a.iterations, a.evaluations) == (b.iterations, b.evaluations)
            np.testing.assert_array_equal(np.array(a.P_av), np.array(b.P_av))
    for x, y in zip(fa, fb):
        for k in x:
            np.testing.assert_array_equal(x[k], y[k], err_msg=k)


C4_FIXTURE = os.path.join(GOLDEN, "c4_fixture.npz")


@pytest.mark.skipif(not os.path.exists(C4_FIXTURE), reason="c4 fixture not generated")
def test_c4_batched_against_oracle(lib):
    """The batched workload's RVEs (generator seed 7, edges 16, 24, 32, 48,
    10-200 grains, tension along random directions, 10 increments of 0.5 s)
    run together through the batched server under a memory budget that
    admits only some of them at once, against the oracle's runs of the same
    RVEs: identical iterations and evaluations in every increment, <P>
    within 1e-10 relative at every increment, P and F on the sampled cells
    after the last increment by tail_close (stopping decisions of the
    material point moved by rounding in a few cells)."""
    fx = np.load(C4_FIXTURE)
    specs = batched_specs(lib, 64, seed=7, **EXAMPLE_NUMERICS)
    chosen = [specs[i] for i in fx["index"]]
    print(f"[c4] edges {[s.cells[0] for s, _ in chosen]}", flush=True)
    fps = [lib.sc_rve_footprint(C.byref(s.desc)) for s, _ in chosen]
    # counts and <P> with about four jobs admitted at once (the others wait
    # for memory; a job's memory is released after its last increment)
    budget = int(sum(sorted(fps)[-4:]))
    srv = Server(lib, budget, len(chosen), batched=True, max_wait_ms=20.0)
    jobs = [srv.add_job(s, [tension_x(t=5.0, N=10, rot_q=q)]) for s, q in chosen]
    reqs = [[srv.submit(j) for _ in range(10)] for j in jobs]
    res = [[srv.wait(r, 600.0) for r in rs] for rs in reqs]
    st = srv.stats()
    srv.close()
    assert st.peak_bytes <= budget and st.max_batch >= 2
    # fields: all jobs resident, kept after their last increment
    srv = Server(lib, 1 << 38, len(chosen), batched=True, max_wait_ms=20.0)
    jobs = [srv.add_job(s, [tension_x(t=5.0, N=10, rot_q=q)], keep_resident=True) for s, q in chosen]
    reqs = [[srv.submit(j) for _ in range(10)] for j in jobs]
    res2 = [[srv.wait(r, 600.0) for r in rs] for rs in reqs]
    fields = [{k: srv.field(j, k, s.N) for k in ("F", "P")} for j, (s, _) in zip(jobs, chosen)]
    for j in jobs:
        srv.remove_job(j)
    srv.close()
    for a_job, b_job in zip(res, res2):  # any admission order, the same results
        for a, b in zip(a_job, b_job):
            np.testing.assert_array_equal(np.array(a.P_av), np.array(b.P_av))
    for j, (spec, _) in enumerate(chosen):
        its = np.array([r.iterations for r in res[j]])
        evs = np.array([r.evaluations for r in res[j]])
        assert all(r.status == 0 for r in res[j])
        np.testing.assert_array_equal(its, fx[f"iterations_{j}"], err_msg=f"rve {j} iterations")
        np.testing.assert_array_equal(evs, fx[f"evaluations_{j}"], err_msg=f"rve {j} evaluations")
        pav = np.array([r.P_av for r in res[j]])
        ref = fx[f"P_av_{j}"]
        assert np.abs(pav - ref).max() / np.abs(ref).max() < 1e-10, f"rve {j} <P>"
        cells = fx[f"cells_{j}"]
        for k in ("P", "F"):
            got = fields[j][k].reshape(spec.N, 9)[cells]
            tail_close(got, fx[f"{k}_{j}"], len(cells), f"rve {j} {k} after 10 increments")


C5_FIXTURE = os.path.join(GOLDEN, "c5_fixture.npz")


def test_serving_stream_alone_equals_mix(lib):
    """The serving workload's first 24 requests (seed 11: Poisson arrivals at
    64 per second, edges log-uniform in [16, 128] rounded to 2^a 3^b -- here
    16, 18, 24, 27, 32, 48, 54, 64, 72, 81, 96, 108 and 128 -- one increment
    of 1e-4 strain each) submitted at their arrival times to the batched
    server, against each request run alone on an idle server: the same
    iterations and evaluations and bitwise the same <P> and P.  Requests on
    grids of at most 36^3 cells also match the oracle's runs
    (c5_fixture.npz): identical counts, <P> and P on sampled cells within
    1e-10."""
    import time

    from solver_cpp_amd.serving import serving_requests

    reqs = serving_requests(lib, 24, rate=64.0, seed=11, edge_range=(16, 128), **EXAMPLE_NUMERICS)
    print(f"[serving] edges {[s.cells[0] for _, s in reqs]}", flush=True)
    load = [tension_x(t=0.1, N=1)]

    def fields(srv, job, spec):
        return srv.field(job, "P", spec.N)

    alone = []
    srv = Server(lib, 1 << 38, 1)
    for _, spec in reqs:
        job = srv.add_job(spec, load, keep_resident=True)
        r = srv.wait(srv.submit(job), 300.0)
        alone.append((r, fields(srv, job, spec)))
        srv.remove_job(job)
    srv.close()
    srv = Server(lib, 1 << 38, 16, batched=True, max_wait_ms=5.0)
    t0 = time.perf_counter()
    ids = []
    for arr, spec in reqs:
        delay = arr - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        job = srv.add_job(spec, load, keep_resident=True)
        ids.append((job, srv.submit(job)))
    mix = []
    for (job, rid), (_, spec) in zip(ids, reqs):
        r = srv.wait(rid, 300.0)
        mix.append((r, fields(srv, job, spec)))
    st = srv.stats()
    for job, _ in ids:
        srv.remove_job(job)
    srv.close()
    assert st.max_concurrent >= 2 and st.max_batch >= 2, (st.max_concurrent, st.max_batch)
    for (a, Pa), (m, Pm) in zip(alone, mix):
        assert a.status == 0 and m.status == 0
        assert (a.iterations, a.evaluations) == (m.iterations, m.evaluations)
        np.testing.assert_array_equal(np.array(a.P_av), np.array(m.P_av))
        np.testing.assert_array_equal(Pa, Pm)
    if not os.path.exists(C5_FIXTURE):
        pytest.skip("c5 fixture not generated")
    fx = np.load(C5_FIXTURE)
    for i in fx["index"]:
        (r, P), spec = mix[i], reqs[i][1]
        assert (r.iterations, r.evaluations) == (int(fx[f"iterations_{i}"]), int(fx[f"evaluations_{i}"])), i
        ref = fx[f"P_av_{i}"]
        assert np.abs(np.array(r.P_av) - ref).max() / np.abs(ref).max() < 1e-10, i
        cells = fx[f"cells_{i}"]
        want = fx[f"P_{i}"]
        err = np.abs(P.reshape(spec.N, 9)[cells] - want).max() / np.abs(want).max()
        assert err < 1e-10, (i, err)
