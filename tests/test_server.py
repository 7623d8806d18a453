"""The job server: continuous batching of RVE increments with HBM-footprint
admission.  A result must be the same alone and in any mix, and the same as
the single-context path's."""
import ctypes as C

import numpy as np
import pytest

from conftest import PRODUCT_SO
from solver_cpp_amd.capi import Library
from solver_cpp_amd.geometry import polycrystal
from solver_cpp_amd.rve import Rve, RveSpec, tension_x
from solver_cpp_amd.serving import Server, batched_specs, rotation_to


@pytest.fixture(scope="module")
def lib():
    return Library(PRODUCT_SO, boundary=True)


def _spec(lib, edge, grains, seed, **kw):
    mat, ori, size = polycrystal((edge, edge, edge), grains, seed=seed)
    return RveSpec(lib, (edge,) * 3, size, mat, ori, itmin=4, itmax=40, **kw)


def test_footprint_and_budget(lib):
    """(CPU) Footprints grow with the grid; a job that cannot fit is refused."""
    small, big = _spec(lib, 8, 3, 1), _spec(lib, 16, 5, 2)
    fs, fb = lib.sc_rve_footprint(C.byref(small.desc)), lib.sc_rve_footprint(C.byref(big.desc))
    assert 0 < fs < fb
    srv = Server(lib, fb // 2, 1)
    with pytest.raises(RuntimeError, match="budget"):
        srv.add_job(big, [tension_x(t=1.0, N=2)])
    srv.close()


def test_invalid_job_is_refused_and_holds_no_memory(lib):
    """(CPU) A job that could not be started -- a material index out of range,
    a grid length with a prime factor above 13 -- is refused when it is
    added, with the reason, and leaves the memory accounting at zero."""
    srv = Server(lib, 1 << 34, 2)
    bad = _spec(lib, 8, 3, 1)
    bad._mat[5] = 7  # only 3 materials
    with pytest.raises(RuntimeError, match="material_at out of range"):
        srv.add_job(bad, [tension_x(t=1.0, N=2)])
    mat, ori, size = polycrystal((17, 8, 8), 3, seed=2)
    odd = RveSpec(lib, (17, 8, 8), size, mat, ori)
    with pytest.raises(RuntimeError, match="prime factor"):
        srv.add_job(odd, [tension_x(t=1.0, N=2)])
    st = srv.stats()
    assert st.used_bytes == 0 and st.resident_jobs == 0 and st.peak_bytes == 0
    srv.close()


def test_rotation_to():
    for u in ([0, 0, 1], [1, 1, 1], [-1, 0.2, 0.1]):
        q = np.array(rotation_to(u))
        w, v = q[0], q[1:]
        x = np.array([1.0, 0, 0])
        rx = x + 2 * w * np.cross(v, x) + 2 * np.cross(v, np.cross(v, x))
        np.testing.assert_allclose(rx, np.array(u) / np.linalg.norm(u), atol=1e-14)


@pytest.mark.gpu
def test_server_matches_single_context(lib):
    """Three increments of one job on the server equal three increments of
    the same RVE driven directly (identical code path and arithmetic)."""
    spec = _spec(lib, 12, 6, 3)
    load = [tension_x(t=1.5, N=3)]
    srv = Server(lib, 1 << 36, 2)
    job = srv.add_job(spec, load, keep_resident=True)
    res = [srv.wait(srv.submit(job)) for _ in range(3)]
    srv.close()
    mat, ori, size = polycrystal((12,) * 3, 6, seed=3)
    r = Rve(lib, (12,) * 3, size, mat, ori, itmin=4, itmax=40)
    r.init()
    for k, inc in enumerate(r.increments(load)):
        assert res[k].status == 0
        assert res[k].iterations == inc["iterations"] and res[k].evaluations == inc["evaluations"]
        if k == 2:
            np.testing.assert_array_equal(np.array(res[k].P_av), np.array(r.scalars().P_av))


@pytest.mark.gpu
def test_alone_equals_mix_and_admission(lib):
    """Requests run alone (one worker, one at a time) and mixed (four workers,
    budget for two resident jobs, requests interleaved) give bitwise equal
    results; the budget is never exceeded and work does overlap."""
    specs = batched_specs(lib, 6, seed=5, edges=(8, 12, 16), grains=(3, 12), itmin=4, itmax=40)
    loads = [[tension_x(t=0.6, N=3, rot_q=q)] for _, q in specs]
    alone = []
    srv = Server(lib, 1 << 36, 1)
    for (spec, _), load in zip(specs, loads):
        job = srv.add_job(spec, load, keep_resident=True)
        alone.append([srv.wait(srv.submit(job)) for _ in range(3)])
        srv.remove_job(job)
    srv.close()
    fp = max(lib.sc_rve_footprint(C.byref(s.desc)) for s, _ in specs)
    budget = 2 * fp
    srv = Server(lib, budget, 4)
    jobs = [srv.add_job(spec, load, keep_resident=True) for (spec, _), load in zip(specs, loads)]
    reqs = [[] for _ in jobs]
    for k in range(3):  # interleave the jobs' increments
        for j, job in enumerate(jobs):
            reqs[j].append(srv.submit(job))
    for job in jobs:  # released as soon as their last increment is done
        srv.remove_job(job)
    mixed = [[srv.wait(r, 100.0) for r in rs] for rs in reqs]
    st = srv.stats()
    assert st.used_bytes == 0 and st.resident_jobs == 0
    srv.close()
    assert st.peak_bytes <= budget
    assert st.max_concurrent >= 2
    for a_job, m_job in zip(alone, mixed):
        for a, m in zip(a_job, m_job):
            assert a.status == 0 and m.status == 0
            assert (a.iterations, a.evaluations, a.increment) == (m.iterations, m.evaluations, m.increment)
            np.testing.assert_array_equal(np.array(a.P_av), np.array(m.P_av))


@pytest.mark.gpu
def test_checkpoint_restart_continues_bitwise(lib, tmp_path):
    """A context checkpointed after three increments and restored into a new
    context from the same descriptor continues exactly: the next increment
    takes the same iterations and gives bitwise the same fields."""
    mat, ori, size = polycrystal((12, 12, 12), 6, seed=3)
    load = tension_x(t=3.0, N=6)
    a = Rve(lib, (12, 12, 12), size, mat, ori, itmin=4, itmax=40)
    a.init()
    gen = a.increments([load])
    for _ in range(3):
        next(gen)
    path = str(tmp_path / "rve.ckpt")
    a.checkpoint(path)
    b = Rve(lib, (12, 12, 12), size, mat, ori, itmin=4, itmax=40)
    b.restore(path)
    out = []
    for r in (a, b):  # increment 4 of the same load case, driven identically
        r.forward(False, True, 0.5, 0.5, 1.5, load)
        sol = r.calculate_solution()
        r.update_coords()
        out.append((sol.required_iterations, sol.converged, {k: r.field(k) for k in ("F", "P", "Fp", "xi", "x_p")}))
    (ia, ca, fa), (ib, cb, fb) = out
    assert ia == ib and ca and cb
    for k in fa:
        np.testing.assert_array_equal(fa[k], fb[k], err_msg=k)
    b.close()
    a.close()


@pytest.mark.gpu
def test_restored_run_matches_uninterrupted_oracle(lib, tmp_path):
    """A context checkpointed after three increments of the tension case,
    restored into a new context and driven through two more increments
    matches the oracle's uninterrupted five increments: identical iterations
    and evaluations, F, P, Fp and xi within 1e-10 (the restart of
    phase_mechanical.f90:1264-1298 and the grid solver's F, F_lastInc)."""
    from conftest import ORACLE_SO

    oracle = Library(ORACLE_SO)
    mat, ori, size = polycrystal((12, 12, 12), 6, seed=3)
    load = [tension_x(t=2.5, N=5)]
    a = Rve(lib, (12, 12, 12), size, mat, ori, itmin=4, itmax=40)
    a.init()
    gen = a.increments(load)
    first = [next(gen) for _ in range(3)]
    path = str(tmp_path / "rve.ckpt")
    a.checkpoint(path)
    a.close()
    b = Rve(lib, (12, 12, 12), size, mat, ori, itmin=4, itmax=40)
    b.restore(path)
    rest = []
    for _ in range(2):  # increments 4 and 5, driven as increments() drives them
        b.forward(False, True, 0.5, 0.5, 2.5 - 0.5 * (3 + len(rest)), load[0])
        sol = b.calculate_solution()
        assert sol.converged
        b.update_coords()
        rest.append((sol.required_iterations, sol.residual_evaluations))
    o = Rve(oracle, (12, 12, 12), size, mat, ori, itmin=4, itmax=40)
    o.init()
    ref = [(d["iterations"], d["evaluations"]) for d in o.increments(load)]
    assert [(d["iterations"], d["evaluations"]) for d in first] + rest == ref
    for k in ("F", "P", "Fp", "xi"):
        x, y = b.field(k), o.field(k)
        assert np.abs(x - y).max() / np.abs(y).max() <= 1e-10, k
    b.close()
    o.close()


@pytest.mark.gpu
def test_eviction_to_host_changes_no_result(lib):
    """Four kept-resident jobs advanced in rounds with HBM for two: idle jobs
    are checkpointed to host memory and restored on their next request, and
    every result is bitwise the one of the job run alone."""
    specs = [_spec(lib, 12, 4 + j, 20 + j) for j in range(4)]
    loads = [[tension_x(t=0.6, N=3)] for _ in specs]
    alone = []
    srv = Server(lib, 1 << 36, 1)
    for spec, load in zip(specs, loads):
        job = srv.add_job(spec, load, keep_resident=True)
        alone.append([srv.wait(srv.submit(job)) for _ in range(3)])
        srv.remove_job(job)
    srv.close()
    fp = max(lib.sc_rve_footprint(C.byref(s.desc)) for s in specs)
    srv = Server(lib, 2 * fp, 2)
    jobs = [srv.add_job(spec, load, keep_resident=True) for spec, load in zip(specs, loads)]
    mixed = [[] for _ in jobs]
    for k in range(3):  # one increment of every job per round; jobs are idle between rounds
        reqs = [srv.submit(job) for job in jobs]
        for j, r in enumerate(reqs):
            mixed[j].append(srv.wait(r, 100.0))
    st = srv.stats()
    for job in jobs:
        srv.remove_job(job)
    srv.close()
    assert st.evictions > 0 and st.peak_bytes <= 2 * fp
    for a_job, m_job in zip(alone, mixed):
        for a, m in zip(a_job, m_job):
            assert a.status == 0 and m.status == 0
            assert a.iterations == m.iterations and a.evaluations == m.evaluations
            np.testing.assert_array_equal(np.array(a.P_av), np.array(m.P_av))
