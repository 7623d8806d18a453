"""Host-side mirror of the reference's grid solver interface.

``Rve`` wraps one ``sc_rve`` context of a loaded library and exposes the
entry points of ``MechSolverBasic`` / ``DamaskGrid`` (src/grid/mech_solver_basic.h,
src/grid/damask_grid.h): ``init``, ``forward``, ``calculate_solution``,
``update_coords`` and ``loop_over_loadcase``.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from .capi import (KT_CLASSES, SC_DEF, SC_FIELD, SC_STRESS, GridNumerics, KernelTimes, Library, LoadStep, Material,
                   PhaseNumerics, RveDesc, Solution, SolverScalars, Stats, ptr)


@dataclass
class Load:
    """One load step (Config::LoadStep, src/grid/config.h:86-96).

    ``deformation`` / ``stress`` are 3x3 nested lists whose entries are
    numbers or the string ``'x'`` (not prescribed), as in the load YAML.
    """
    deformation: list
    stress: list
    deformation_type: str = "dot_F"
    stress_type: str = "P"
    t: float = 5.0
    N: int = 10
    r: float = 1.0
    rot_q: Sequence[float] = (1.0, 0.0, 0.0, 0.0)
    estimate_rate: bool = True
    f_out: int = 1

    def to_c(self) -> LoadStep:
        s = LoadStep()
        s.deformation_type = SC_DEF[self.deformation_type]
        s.stress_type = SC_STRESS[self.stress_type]
        for i in range(3):
            for j in range(3):
                q = i + 3 * j  # column-major
                d = self.deformation[i][j]
                p = self.stress[i][j]
                s.deformation_mask[q] = 1 if d == "x" else 0
                s.deformation[q] = 0.0 if d == "x" else float(d)
                s.stress_mask[q] = 1 if p == "x" else 0
                s.stress[q] = 0.0 if p == "x" else float(p)
        for k in range(4):
            s.rot_q[k] = float(self.rot_q[k])
        s.t, s.N, s.r = float(self.t), int(self.N), float(self.r)
        s.estimate_rate = 1 if self.estimate_rate else 0
        s.f_out = int(self.f_out)
        return s


def tension_x(t: float = 5.0, N: int = 10, rate: float = 1.0e-3, rot_q=(1.0, 0.0, 0.0, 0.0)) -> Load:
    """examples/grid/tensionX.yaml: dot_F_xx = rate, P_yy = P_zz = 0."""
    return Load(deformation=[[rate, 0, 0], [0, "x", 0], [0, 0, "x"]],
                stress=[["x", "x", "x"], ["x", 0, "x"], ["x", "x", 0]], t=t, N=N, rot_q=rot_q)


class RveSpec:
    """An sc_rve_desc together with the arrays it points to (kept alive here)."""

    def __init__(self, lib: Library, cells, size, material_at: np.ndarray, orientations: np.ndarray,
                 material: Optional[Material] = None, phase_num: Optional[PhaseNumerics] = None,
                 grid_num: Optional[GridNumerics] = None, precision: int = 0, **grid_overrides):
        m, p, g = lib.defaults()
        if material is not None:
            m = material
        if phase_num is not None:
            p = phase_num
        if grid_num is not None:
            g = grid_num
        for k, v in grid_overrides.items():
            setattr(g, k, v)
        self.cells = tuple(int(c) for c in cells)
        self.N = int(np.prod(self.cells))
        self._mat = np.ascontiguousarray(material_at, dtype=np.int32).reshape(-1)
        self._ori = np.ascontiguousarray(orientations, dtype=np.float64).reshape(-1, 4)
        assert self._mat.size == self.N
        d = RveDesc()
        for k in range(3):
            d.cells[k] = self.cells[k]
            d.size[k] = float(size[k])
        d.n_materials = self._ori.shape[0]
        d.orientations = ptr(self._ori)
        d.material_at = ptr(self._mat, C.POINTER(C.c_int))
        d.material, d.phase_num, d.grid_num = m, p, g
        d.precision = int(precision)
        self.desc = d


class Rve:
    """One representative volume element on one library backend."""

    def __init__(self, lib: Library, cells, size, material_at: np.ndarray, orientations: np.ndarray,
                 material: Optional[Material] = None, phase_num: Optional[PhaseNumerics] = None,
                 grid_num: Optional[GridNumerics] = None, precision: int = 0, **grid_overrides):
        self.lib = lib
        self.spec = RveSpec(lib, cells, size, material_at, orientations, material, phase_num, grid_num, precision,
                            **grid_overrides)
        self.cells, self.N, self.desc = self.spec.cells, self.spec.N, self.spec.desc
        h = C.c_void_p()
        lib.check(lib.sc_rve_create(C.byref(self.desc), C.byref(h)), "sc_rve_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.sc_rve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # MechSolverBasic entry points --------------------------------------------
    def init(self):
        self.lib.check(self.lib.sc_rve_init(self.h), "sc_rve_init")

    def forward(self, cut_back: bool, guess: bool, dt: float, dt_old: float, t_remaining: float, load: Load):
        st = load.to_c()
        self.lib.check(self.lib.sc_rve_forward(self.h, int(cut_back), int(guess), dt, dt_old, t_remaining,
                                               C.byref(st)), "sc_rve_forward")

    def calculate_solution(self) -> Solution:
        s = Solution()
        self.lib.check(self.lib.sc_rve_solution(self.h, C.byref(s)), "sc_rve_solution")
        return s

    def update_coords(self):
        self.lib.check(self.lib.sc_rve_update_coords(self.h), "sc_rve_update_coords")

    def loop_over_loadcase(self, loads: Sequence[Load]):
        """DamaskGrid::loop_over_loadcase; returns per-increment (iterations, evaluations)."""
        steps = (LoadStep * len(loads))(*[l.to_c() for l in loads])
        n = sum(int(l.N) for l in loads)
        its = np.zeros(n, dtype=np.int32)
        evs = np.zeros(n, dtype=np.int32)
        rc = self.lib.sc_rve_run_loadcase(self.h, steps, len(loads), ptr(its, C.POINTER(C.c_int)),
                                          ptr(evs, C.POINTER(C.c_int)), n)
        self.lib.check(rc, "sc_rve_run_loadcase")
        return its, evs

    def form_residual(self, F: np.ndarray) -> np.ndarray:
        F = np.ascontiguousarray(F, dtype=np.float64).reshape(-1)
        r = np.zeros_like(F)
        self.lib.check(self.lib.sc_rve_form_residual(self.h, ptr(F), ptr(r)), "sc_rve_form_residual")
        return r

    # restart --------------------------------------------------------------------
    def checkpoint(self, path: str):
        """Write the state between increments to `path` (sc_rve_checkpoint)."""
        self.lib.check(self.lib.sc_rve_checkpoint(self.h, os.fsencode(path)), "sc_rve_checkpoint")

    def restore(self, path: str):
        """Read a checkpoint of a context built from the same descriptor."""
        self.lib.check(self.lib.sc_rve_restore(self.h, os.fsencode(path)), "sc_rve_restore")

    # DADF5 results -----------------------------------------------------------------
    def result_open(self, path: str, material_yaml: str = None, numerics_yaml: str = None,
                    compression: bool = True):
        """Create the DADF5 job file (sc_rve_result_open); run_loadcase then
        writes increments at each load step's f_out cadence."""
        from .capi import ResultDesc

        d = ResultDesc(material_yaml.encode() if material_yaml is not None else None,
                       numerics_yaml.encode() if numerics_yaml is not None else None, int(compression))
        self._result_desc = d
        self.lib.check(self.lib.sc_rve_result_open(self.h, os.fsencode(path), C.byref(d)), "sc_rve_result_open")

    def result_write(self, inc: int, t: float):
        """Write increment_<inc> of the open job file (materialpoint_result)."""
        self.lib.check(self.lib.sc_rve_result_write(self.h, int(inc), float(t)), "sc_rve_result_write")

    def result_close(self):
        self.lib.check(self.lib.sc_rve_result_close(self.h), "sc_rve_result_close")

    # access -------------------------------------------------------------------
    def field(self, name: str) -> np.ndarray:
        n = {"xi": 12, "gamma": 12, "x_p": 3, "iterations": 2}.get(name, 9) * self.N
        if name == "x_n":
            n = 3 * (self.cells[0] + 1) * (self.cells[1] + 1) * (self.cells[2] + 1)
        out = np.zeros(n)
        self.lib.check(self.lib.sc_rve_get_field(self.h, SC_FIELD[name], ptr(out), n), f"get_field {name}")
        return out

    def set_field(self, name: str, values: np.ndarray):
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        self.lib.check(self.lib.sc_rve_set_field(self.h, SC_FIELD[name], ptr(v), v.size), f"set_field {name}")

    def scalars(self) -> SolverScalars:
        s = SolverScalars()
        self.lib.check(self.lib.sc_rve_get_scalars(self.h, C.byref(s)), "get_scalars")
        return s

    def stats(self) -> Stats:
        s = Stats()
        self.lib.check(self.lib.sc_rve_get_stats(self.h, C.byref(s)), "get_stats")
        return s

    def reset_stats(self):
        self.lib.check(self.lib.sc_rve_reset_stats(self.h), "reset_stats")

    # measurement (HIP library) -------------------------------------------------
    def profile(self, enable: bool = True, classes: Optional[Sequence[str]] = None):
        """Kernel timing with HIP events: every class, or only the named ones
        (names of KT_CLASSES)."""
        if enable and classes is not None:
            mask = sum(1 << KT_CLASSES.index(c) for c in classes)
            self.lib.check(self.lib.sc_rve_profile_classes(self.h, mask), "profile")
        else:
            self.lib.check(self.lib.sc_rve_profile(self.h, int(enable)), "profile")

    def kernel_times(self) -> dict:
        k = KernelTimes()
        self.lib.check(self.lib.sc_rve_get_kernel_times(self.h, C.byref(k)), "kernel_times")
        return {name: dict(launches=int(k.launches[i]), ms=float(k.ms[i]), points=int(k.points[i]))
                for i, name in enumerate(KT_CLASSES)}

    # increment loop -------------------------------------------------------------
    def increments(self, loads: Sequence[Load], max_cut_back: Optional[int] = None):
        """DamaskGrid::loop_over_loadcase (damask_grid.cpp:82-196) driven from
        the host one increment at a time: yields one dict per converged
        increment (index, time, iterations, evaluations, cut-backs, wall
        seconds).  Same control flow as sc_rve_run_loadcase."""
        import time as _time
        if max_cut_back is None:
            max_cut_back = int(self.desc.grid_num.max_cut_back)
        t, dt, dt_prev = 0.0, 1.0, 0.0
        cut_back, cut_back_level, total_inc = False, 0, 0
        for l, load in enumerate(loads):
            t_0 = t
            guess = bool(load.estimate_rate) and l > 0
            for inc in range(int(load.N)):
                total_inc += 1
                w0 = _time.perf_counter()
                dt_prev = dt
                if abs(load.r - 1.0) < 1e-9:
                    dt = load.t / load.N
                else:
                    dt = load.t * (load.r ** inc - load.r ** (inc + 1)) / (1.0 - load.r ** load.N)
                dt *= 2.0 ** (-cut_back_level)
                step_fraction, its, evs, cuts = 0, 0, 0, 0
                while step_fraction < 2 ** cut_back_level:
                    t_remaining = load.t + t_0 - t
                    t += dt
                    step_fraction += 1
                    self.forward(cut_back, guess, dt, dt_prev, t_remaining, load)
                    sol = self.calculate_solution()
                    its += sol.required_iterations
                    evs += sol.residual_evaluations
                    if sol.converged and sol.stag_converged and not sol.terminally_ill:
                        self.update_coords()
                        dt_prev = dt
                        cut_back = False
                        guess = True
                    elif cut_back_level < max_cut_back:
                        cut_back = True
                        step_fraction = (step_fraction - 1) * 2
                        cut_back_level += 1
                        t -= dt
                        dt /= 2.0
                        cuts += 1
                    else:
                        raise RuntimeError(f"no solution found for increment {total_inc}")
                cut_back_level = max(0, int(cut_back) - 1)
                yield dict(increment=total_inc, t=t, iterations=its, evaluations=evs, cutbacks=cuts,
                           seconds=_time.perf_counter() - w0)
