"""ctypes description of include/solver_cpp_amd.h.

The same declarations bind the HIP product library and (in tests only) the
CPU oracle library, which exports the same ``sc_*`` symbols.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Material(C.Structure):
    """sc_material: cF phenopowerlaw phase (examples/grid/material.yaml:7-24)."""
    _fields_ = [
        ("C_11", C.c_double), ("C_12", C.c_double), ("C_44", C.c_double),
        ("xi_0_sl", C.c_double), ("xi_inf_sl", C.c_double), ("h_0_sl_sl", C.c_double),
        ("n_sl", C.c_double), ("a_sl", C.c_double), ("dot_gamma_0_sl", C.c_double),
        ("h_sl_sl", C.c_double * 7),
        ("atol_xi", C.c_double), ("atol_gamma", C.c_double),
    ]


class PhaseNumerics(C.Structure):
    """sc_phase_numerics (phase_mechanical.f90:280-294 defaults)."""
    _fields_ = [
        ("r_cutback_min", C.c_double), ("r_cutback", C.c_double), ("r_increase", C.c_double),
        ("eps_rel_state", C.c_double), ("N_iter_state_max", C.c_int), ("N_iter_Lp_max", C.c_int),
        ("r_linesearch_Lp", C.c_double), ("eps_rel_Lp", C.c_double), ("eps_abs_Lp", C.c_double),
    ]


class GridNumerics(C.Structure):
    """sc_grid_numerics (src/grid/config.h:50-78)."""
    _fields_ = [
        ("itmin", C.c_int), ("itmax", C.c_int), ("update_gamma", C.c_int),
        ("eps_abs_div", C.c_double), ("eps_rel_div", C.c_double),
        ("eps_abs_P", C.c_double), ("eps_rel_P", C.c_double),
        ("divergence_correction", C.c_int), ("derivative", C.c_int),
        ("stag_iter_max", C.c_int), ("max_cut_back", C.c_int), ("ngmres_m", C.c_int),
        ("fft_plan_mode", C.c_int),
    ]


class RveDesc(C.Structure):
    _fields_ = [
        ("cells", C.c_int * 3), ("size", C.c_double * 3), ("n_materials", C.c_int),
        ("orientations", _dp), ("material_at", _ip),
        ("material", Material), ("phase_num", PhaseNumerics), ("grid_num", GridNumerics),
        ("precision", C.c_int),
    ]


class LoadStep(C.Structure):
    _fields_ = [
        ("deformation_type", C.c_int), ("deformation", C.c_double * 9), ("deformation_mask", C.c_int * 9),
        ("stress_type", C.c_int), ("stress", C.c_double * 9), ("stress_mask", C.c_int * 9),
        ("rot_q", C.c_double * 4), ("t", C.c_double), ("N", C.c_int), ("r", C.c_double),
        ("estimate_rate", C.c_int), ("f_out", C.c_int),
    ]


class Solution(C.Structure):
    _fields_ = [
        ("converged", C.c_int), ("stag_converged", C.c_int), ("terminally_ill", C.c_int),
        ("required_iterations", C.c_int), ("residual_evaluations", C.c_int),
        ("err_div", C.c_double), ("err_BC", C.c_double),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("residual_evaluations", C.c_int64), ("material_point_updates", C.c_int64),
        ("increments", C.c_int64), ("cutbacks", C.c_int64),
        ("newton_iterations", C.c_int64), ("state_iterations", C.c_int64),
        ("seconds_total", C.c_double), ("fp32_evaluations", C.c_int64),
    ]


class SolverScalars(C.Structure):
    _fields_ = [
        ("P_av", C.c_double * 9), ("P_aim", C.c_double * 9), ("F_aim", C.c_double * 9),
        ("F_aim_dot", C.c_double * 9), ("F_aim_last_inc", C.c_double * 9),
        ("C_volAvg", C.c_double * 81), ("C_minMaxAvg", C.c_double * 81),
        ("S", C.c_double * 81), ("C_ref", C.c_double * 81),
        ("err_div", C.c_double), ("err_BC", C.c_double), ("total_iter", C.c_int),
    ]


class RequestResult(C.Structure):
    _fields_ = [
        ("request", C.c_int64), ("job", C.c_int64), ("status", C.c_int), ("increment", C.c_int),
        ("iterations", C.c_int), ("evaluations", C.c_int), ("P_av", C.c_double * 9), ("t", C.c_double),
        ("material_point_updates", C.c_int64),
        ("t_submit", C.c_double), ("t_start", C.c_double), ("t_end", C.c_double),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("budget_bytes", C.c_int64), ("used_bytes", C.c_int64), ("peak_bytes", C.c_int64),
        ("resident_jobs", C.c_int64), ("completed_requests", C.c_int64), ("max_concurrent", C.c_int),
        ("evictions", C.c_int64), ("batches", C.c_int64), ("batched_evaluations", C.c_int64),
        ("max_batch", C.c_int64), ("batch_seconds", C.c_double),
    ]


class MpDevicePlanes(C.Structure):
    """sc_mp_device_planes: device addresses of the structure-of-arrays planes
    of sc_mp_update_noalloc (0 for an optional array left out)."""
    _fields_ = [(k, C.c_void_p) for k in ("F", "F0", "Fp0", "Lp0", "S0", "xi0", "gamma0", "Lp", "P", "Fp", "S",
                                          "Fe", "xi", "gamma", "dPdF", "broken", "iterations")]


KT_CLASSES = ["response", "tangent", "fft_fwd", "fft_bwd", "gamma", "reduce", "other"]


class KernelTimes(C.Structure):
    _fields_ = [("launches", C.c_int64 * 7), ("ms", C.c_double * 7), ("points", C.c_int64 * 7)]


SC_FIELD = dict(F=0, P=1, Fp=2, Lp=3, S=4, Fe=5, F_last_inc=6, xi=7, gamma=8, x_p=9, x_n=10, iterations=11)
SC_DEF = dict(F=0, dot_F=1, L=2)
SC_STRESS = dict(P=0, dot_P=1)
PRECISION = dict(fp64=0, mixed=1)

_SIGNATURES = {
    "sc_last_error": (C.c_char_p, []),
    "sc_backend_name": (C.c_char_p, []),
    "sc_default_material": (None, [C.POINTER(Material)]),
    "sc_default_phase_numerics": (None, [C.POINTER(PhaseNumerics)]),
    "sc_default_grid_numerics": (None, [C.POINTER(GridNumerics)]),
    "sc_rve_create": (C.c_int, [C.POINTER(RveDesc), C.POINTER(C.c_void_p)]),
    "sc_rve_destroy": (None, [C.c_void_p]),
    "sc_rve_init": (C.c_int, [C.c_void_p]),
    "sc_rve_forward": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                 C.POINTER(LoadStep)]),
    "sc_rve_solution": (C.c_int, [C.c_void_p, C.POINTER(Solution)]),
    "sc_rve_update_coords": (C.c_int, [C.c_void_p]),
    "sc_rve_run_loadcase": (C.c_int, [C.c_void_p, C.POINTER(LoadStep), C.c_int, _ip, _ip, C.c_int]),
    "sc_rve_form_residual": (C.c_int, [C.c_void_p, _dp, _dp]),
    "sc_ngmres_replay": (C.c_int, [C.c_int64, C.c_int, _dp, C.c_int, _dp, _dp]),
    "sc_rve_get_field": (C.c_int, [C.c_void_p, C.c_int, _dp, C.c_int64]),
    "sc_rve_set_field": (C.c_int, [C.c_void_p, C.c_int, _dp, C.c_int64]),
    "sc_rve_get_scalars": (C.c_int, [C.c_void_p, C.POINTER(SolverScalars)]),
    "sc_rve_get_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "sc_rve_reset_stats": (C.c_int, [C.c_void_p]),
    "sc_spectral_xi": (C.c_int, [_ip, _dp, C.c_int, _dp, _dp]),
    "sc_fft_r2c": (C.c_int, [_ip, C.c_int, _dp, _dp]),
    "sc_fft_c2r": (C.c_int, [_ip, C.c_int, _dp, C.c_double, _dp]),
    "sc_divergence_rms": (C.c_int, [_ip, _dp, _dp, _dp, _dp, _dp]),
    "sc_gamma_convolution": (C.c_int, [_ip, _dp, _dp, _dp, _dp]),
    "sc_masked_compliance": (C.c_int, [_dp, _dp, _ip, _dp]),
    "sc_scaled_geom_size": (C.c_int, [_ip, _dp, C.c_int, _dp]),
    "sc_forward_field": (C.c_int, [C.c_int64, C.c_double, _dp, _dp, _dp, _dp]),
    "sc_calculate_rate": (C.c_int, [C.c_int64, C.c_int, _dp, _dp, C.c_double, _dp, _dp]),
    "sc_response_reductions": (C.c_int, [C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp]),
    "sc_mp_update": (C.c_int, [C.c_int64, C.c_double, C.POINTER(Material), C.POINTER(PhaseNumerics)]
                     + [_dp] * 7 + [_dp] * 7 + [_dp, _ip, _ip]),
    "sc_qu2om": (None, [_dp, _dp]),
    "sc_rotational_part": (None, [_dp, _dp]),
    "f_math_invert": (None, [_dp, _ip, _dp, _ip]),
    "f_math_3333to99": (None, [_dp, _dp]),
    "f_math_99to3333": (None, [_dp, _dp]),
    "f_math_mul3333xx33": (None, [_dp, _dp, _dp]),
    "f_rotate_tensor2": (None, [_dp, _dp, _dp, _ip]),
    "f_rotate_tensor4": (None, [_dp, _dp, _dp, _ip]),
}

# Entry points of the HIP library that the oracle does not have, bound when
# present: measurement hooks and restart checkpoints.
PROFILE_SIGNATURES = {
    "sc_rve_profile": (C.c_int, [C.c_void_p, C.c_int]),
    "sc_rve_profile_classes": (C.c_int, [C.c_void_p, C.c_uint]),
    "sc_rve_get_kernel_times": (C.c_int, [C.c_void_p, C.POINTER(KernelTimes)]),
    "sc_rve_checkpoint": (C.c_int, [C.c_void_p, C.c_char_p]),
    "sc_rve_restore": (C.c_int, [C.c_void_p, C.c_char_p]),
    "sc_rve_result_open": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p]),
    "sc_rve_result_write": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "sc_rve_result_close": (C.c_int, [C.c_void_p]),
    "sc_om2qu": (None, [_dp, _dp]),
    "sc_dadf5_create": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_void_p)]),
    "sc_dadf5_dataset_f64": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.POINTER(C.c_int64), _dp]),
    "sc_dadf5_dataset_str": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p]),
    "sc_dadf5_attr_str": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_char_p]),
    "sc_dadf5_attr_f64": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_double]),
    "sc_dadf5_close": (C.c_int, [C.c_void_p]),
    "sc_mp_update_noalloc": (C.c_int, [C.c_int64, C.c_double, C.POINTER(Material), C.POINTER(PhaseNumerics),
                                       C.POINTER(MpDevicePlanes), C.c_void_p]),
}


class ResultDesc(C.Structure):
    """sc_result_desc: DADF5 job file set-up (result.f90:76-116)."""

    _fields_ = [("material_yaml", C.c_char_p), ("numerics_yaml", C.c_char_p), ("compression", C.c_int)]

# Symbols only the product exports: the drop-in boundary of the Fortran core
# (f_*), the job-input readers and the job server.
BOUNDARY_SIGNATURES = {
    "f_materialpoint_initBase": (None, [C.c_char_p, C.c_int, C.c_char_p, C.c_int, C.c_char_p, C.c_int,
                                        C.c_char_p, C.c_int]),
    "f_materialpoint_initDamask": (None, []),
    "f_materialpoint_forward": (None, []),
    "f_materialpoint_result": (None, [_ip, _dp]),
    "f_homogenization_fetch_tensor_pointers": (None, [C.POINTER(_dp)] * 4 + [C.POINTER(C.c_void_p)]),
    "f_homogenization_mechanical_response": (None, [_dp, _ip, _ip]),
    "f_homogenization_thermal_response": (None, [_dp, _ip, _ip]),
    "f_homogenization_mechanical_response2": (None, [_dp, _ip, _ip]),
    "f_discretization_init": (None, [_ip, _ip, _dp, _ip, _dp, _ip, _ip]),
    "f_discretization_fetch_ip_node_coord_pointers": (None, [C.POINTER(_dp), C.POINTER(_dp)]),
    "f_VTI_readCellsSizeOrigin": (None, [C.c_char_p, C.c_int, _ip, _dp, _dp]),
    "f_VTI_readDataset_int": (None, [C.c_char_p, C.c_int, C.c_char_p, C.c_int, _ip, _ip]),
    "sc_boundary_get_state": (C.c_int, [C.c_int, _dp, C.c_int64]),
    "sc_boundary_result_count": (C.c_int, []),
    "sc_boundary_reset": (None, []),
    "sc_boundary_set_result_file": (C.c_int, [C.c_char_p, C.c_int]),
    "sc_read_vti": (C.c_int, [C.c_char_p, _ip, _dp, _dp, _ip, C.c_int64]),
    "sc_read_material": (C.c_int, [C.c_char_p, C.POINTER(Material), _dp, C.c_int, _ip]),
    "sc_read_numerics": (C.c_int, [C.c_char_p, C.POINTER(GridNumerics), C.POINTER(PhaseNumerics)]),
    "sc_read_load": (C.c_int, [C.c_char_p, C.POINTER(LoadStep), C.c_int, _ip]),
    "sc_rve_footprint": (C.c_int64, [C.POINTER(RveDesc)]),
    "sc_server_create": (C.c_int, [C.c_int64, C.c_int, C.POINTER(C.c_void_p)]),
    "sc_server_create_batched": (C.c_int, [C.c_int64, C.c_int, C.c_double, C.POINTER(C.c_void_p)]),
    "sc_server_destroy": (None, [C.c_void_p]),
    "sc_server_add_job": (C.c_int, [C.c_void_p, C.POINTER(RveDesc), C.POINTER(LoadStep), C.c_int, C.c_int,
                                    C.POINTER(C.c_int64)]),
    "sc_server_submit": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]),
    "sc_server_wait": (C.c_int, [C.c_void_p, C.c_int64, C.c_double, C.POINTER(RequestResult)]),
    "sc_server_remove_job": (C.c_int, [C.c_void_p, C.c_int64]),
    "sc_server_get_field": (C.c_int, [C.c_void_p, C.c_int64, C.c_int, _dp, C.c_int64]),
    "sc_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
}


def ptr(a: np.ndarray, kind=_dp):
    """Pointer to a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]
    return a.ctypes.data_as(kind)


class Library:
    """A loaded solver library (product or oracle) with typed entry points."""

    def __init__(self, path: str, boundary: bool = False):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.path = path
        self.lib = C.CDLL(path, mode=C.RTLD_LOCAL)
        sigs = dict(_SIGNATURES)
        if boundary:
            sigs.update(BOUNDARY_SIGNATURES)
        for name, (res, args) in sigs.items():
            fn = getattr(self.lib, name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)
        for name, (res, args) in PROFILE_SIGNATURES.items():
            if hasattr(self.lib, name):
                fn = getattr(self.lib, name)
                fn.restype = res
                fn.argtypes = args
                setattr(self, name, fn)

    def check(self, rc: int, what: str = ""):
        if rc != 0:
            msg = self.sc_last_error()
            raise RuntimeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    @property
    def backend(self) -> str:
        return self.sc_backend_name().decode()

    def defaults(self):
        m, p, g = Material(), PhaseNumerics(), GridNumerics()
        self.sc_default_material(C.byref(m))
        self.sc_default_phase_numerics(C.byref(p))
        self.sc_default_grid_numerics(C.byref(g))
        return m, p, g
