"""ctypes binding of the C ABI in include/aspher_hip.h (libaspher_hip.so).

The library is built in-tree by ``python -m aspher_amd.build``. Loading fails
loudly when it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libaspher_hip.so")

OK, ERR_INVALID, ERR_LOGIC, ERR_RUNTIME, ERR_DEVICE = 0, 1, 2, 3, 4

# Every symbol include/aspher_hip.h declares (checked by the CPU tests).
EXPORTS = (
    "aspher_last_error", "aspher_version", "aspher_device_count",
    "aspher_solver_create", "aspher_solver_destroy", "aspher_solver_matvec",
    "aspher_solver_matvec_device", "aspher_solver_matvec_device_noalloc",
    "aspher_solver_time_matvec", "aspher_solver_time_stages",
    "aspher_solver_solve",
    "aspher_solver_solve_device", "aspher_solver_info", "aspher_nested_params_default",
    "aspher_solve_nested", "aspher_solve_nested_device", "aspher_result_get_info",
    "aspher_result_get_field", "aspher_result_field_device", "aspher_result_get_history",
    "aspher_result_get_stage", "aspher_result_destroy", "aspher_configure_allocator",
    "aspher_love_points", "aspher_love_table", "aspher_stencil_taps",
    "aspher_precond_apply", "aspher_clear_cache", "aspher_device_malloc",
    "aspher_device_free", "aspher_memcpy_htod", "aspher_memcpy_dtoh",
    "aspher_solver_create_shared", "aspher_adjoint_params_default",
    "aspher_solver_adjoint", "aspher_solver_adjoint_device", "aspher_copy_device",
    "aspher_ensemble_params_default", "aspher_solve_ensemble", "aspher_solve_ensemble_device",
    "aspher_rough_surface", "aspher_rough_surface_device",
    "aspher_tangential_create", "aspher_tangential_destroy", "aspher_tangential_matvec",
    "aspher_tangential_matvec_device", "aspher_tangential_precond_apply",
    "aspher_cerruti_points", "aspher_cerruti_tables",
    "aspher_project_load_simplex", "aspher_test_fault",
)


class OperatorInfo(C.Structure):
    _fields_ = [("backend", C.c_char * 16), ("n", C.c_int64), ("padded_side", C.c_int),
                ("bytes", C.c_int64), ("bytes_spectrum", C.c_int64),
                ("bytes_scratch", C.c_int64), ("build_seconds", C.c_double)]


class NestedParams(C.Structure):
    _fields_ = [("coarsest", C.c_int), ("q", C.c_int), ("leaf_side", C.c_int),
                ("precond", C.c_int), ("precond_engine", C.c_char * 16),
                ("precond_radius", C.c_int), ("precond_occupancy_max", C.c_double),
                ("precond_cost_gate", C.c_double), ("coarse_tol", C.c_double),
                ("single_precision", C.c_int), ("precision", C.c_char * 24),
                ("float_floor", C.c_double), ("allow_tolerance_relaxation", C.c_int),
                ("light_result", C.c_int), ("backend", C.c_char * 16),
                ("record_error_history", C.c_int), ("active_set", C.c_int),
                ("active_delta", C.c_double), ("active_halo", C.c_int),
                ("active_max_rounds", C.c_int), ("active_all_levels", C.c_int),
                ("active_occupancy_max", C.c_double), ("reuse_operators", C.c_int),
                ("graph_max_points", C.c_int64)]


class ResultInfo(C.Structure):
    _fields_ = [("grid_size", C.c_int), ("objective", C.c_double),
                ("contact_area", C.c_double), ("mean_pressure", C.c_double),
                ("approach", C.c_double), ("iterations", C.c_int), ("error", C.c_double),
                ("converged", C.c_int), ("active_rounds", C.c_int),
                ("active_fallback", C.c_int), ("status", C.c_char * 32),
                ("status_reason", C.c_char * 32), ("pk_error", C.c_double),
                ("fw_gap", C.c_double), ("fw_error", C.c_double),
                ("load_error", C.c_double), ("pressure_violation", C.c_double),
                ("penetration_error", C.c_double), ("requested_tol", C.c_double),
                ("effective_tol", C.c_double), ("p_ref", C.c_double), ("g_ref", C.c_double),
                ("validation_scope", C.c_char * 32), ("operator_error_kind", C.c_char * 32),
                ("matvec_count", C.c_int64), ("verification_matvec_count", C.c_int64),
                ("precond_count", C.c_int64), ("identification_steps", C.c_int),
                ("returned_best", C.c_int), ("precond_dropped", C.c_int),
                ("mem_cg_state", C.c_int64), ("mem_best_iterate", C.c_int64),
                ("mem_contact_mask", C.c_int64), ("mem_output_arrays", C.c_int64),
                ("mem_peak", C.c_int64), ("t_total", C.c_double), ("t_matvec", C.c_double),
                ("t_precond", C.c_double), ("t_build", C.c_double), ("t_coarse", C.c_double),
                ("t_verification", C.c_double), ("t_candidate", C.c_double),
                ("t_output", C.c_double), ("has_displacement", C.c_int),
                ("has_gap", C.c_int), ("n_history", C.c_int), ("n_stages", C.c_int)]


class StageInfo(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("precision", C.c_char * 16), ("q", C.c_int),
                ("requested_tol", C.c_double), ("effective_tol", C.c_double),
                ("iterations", C.c_int), ("matvec_count", C.c_int64),
                ("seconds", C.c_double), ("status", C.c_char * 32),
                ("fw_error", C.c_double), ("penetration_error", C.c_double),
                ("precond_dropped", C.c_int)]


class AdjointParams(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iter", C.c_int), ("precond", C.c_char * 16)]


class AdjointInfo(C.Structure):
    _fields_ = [("converged", C.c_int), ("iterations", C.c_int), ("n_contact", C.c_int64),
                ("matvec_count", C.c_int64), ("precond_count", C.c_int64),
                ("rel_residual", C.c_double), ("multiplier", C.c_double),
                ("grad_p_nominal", C.c_double), ("seconds", C.c_double),
                ("status", C.c_char * 32), ("precond", C.c_char * 16)]


class EnsembleParams(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iter", C.c_int), ("use_pr", C.c_int),
                ("precision", C.c_char * 16), ("precond", C.c_char * 16),
                ("check_every", C.c_int), ("use_graph", C.c_int), ("autotune", C.c_int),
                ("keep_best", C.c_int), ("stall_limit", C.c_int), ("light_result", C.c_int)]


class EnsembleStats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("setup_seconds", C.c_double),
                ("loop_seconds", C.c_double), ("finish_seconds", C.c_double),
                ("tune_seconds", C.c_double), ("groups", C.c_int), ("replays", C.c_int),
                ("polls", C.c_int), ("host_fallbacks", C.c_int),
                ("device_iterations", C.c_int), ("graph", C.c_int)]


_lib = None


def _declare(L):
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    d, i, s = C.c_double, C.c_int, C.c_char_p
    sig = {
        "aspher_last_error": (s, []),
        "aspher_version": (s, []),
        "aspher_device_count": (i, [ip]),
        "aspher_solver_create": (i, [i, d, d, s, C.POINTER(vp)]),
        "aspher_solver_destroy": (None, [vp]),
        "aspher_solver_matvec": (i, [vp, dp, dp]),
        "aspher_solver_matvec_device": (i, [vp, vp, vp, i]),
        "aspher_solver_matvec_device_noalloc": (i, [vp, vp, vp, i]),
        "aspher_solver_time_matvec": (i, [vp, vp, vp, i, i, dp]),
        "aspher_solver_time_stages": (i, [vp, vp, vp, i, i, dp]),
        "aspher_solver_solve": (i, [vp, dp, d, d, i, i, s, dp, C.POINTER(vp)]),
        "aspher_solver_solve_device": (i, [vp, vp, d, d, i, i, s, vp, C.POINTER(vp)]),
        "aspher_solver_info": (i, [vp, C.POINTER(OperatorInfo)]),
        "aspher_nested_params_default": (None, [C.POINTER(NestedParams)]),
        "aspher_solve_nested": (i, [i, dp, d, d, d, d, i, i, C.POINTER(NestedParams),
                                    C.POINTER(vp)]),
        "aspher_solve_nested_device": (i, [i, vp, d, d, d, d, i, i, C.POINTER(NestedParams),
                                           C.POINTER(vp)]),
        "aspher_result_get_info": (i, [vp, C.POINTER(ResultInfo)]),
        "aspher_result_get_field": (i, [vp, i, dp]),
        "aspher_result_field_device": (i, [vp, i, C.POINTER(vp)]),
        "aspher_result_get_history": (i, [vp, dp]),
        "aspher_result_get_stage": (i, [vp, i, C.POINTER(StageInfo)]),
        "aspher_result_destroy": (None, [vp]),
        "aspher_configure_allocator": (i, [C.c_int64, C.c_int64, ip]),
        "aspher_love_points": (i, [i, dp, dp, dp, dp, dp]),
        "aspher_love_table": (i, [i, d, d, dp]),
        "aspher_stencil_taps": (i, [i, i, ip, ip, ip, dp]),
        "aspher_precond_apply": (i, [i, s, i, i, dp, C.POINTER(C.c_uint8), dp]),
        "aspher_clear_cache": (i, []),
        "aspher_project_load_simplex": (i, [C.c_int64, dp, d, i, dp, dp]),
        "aspher_test_fault": (i, [i, C.c_int64]),
        "aspher_device_malloc": (i, [C.c_int64, C.POINTER(vp)]),
        "aspher_device_free": (i, [vp]),
        "aspher_memcpy_htod": (i, [vp, vp, C.c_int64]),
        "aspher_memcpy_dtoh": (i, [vp, vp, C.c_int64]),
        "aspher_solver_create_shared": (i, [i, d, d, s, C.POINTER(vp)]),
        "aspher_copy_device": (i, [vp, vp, C.c_int64]),
        "aspher_ensemble_params_default": (None, [C.POINTER(EnsembleParams)]),
        "aspher_rough_surface": (i, [i, d, d, d, d, C.c_uint64, dp]),
        "aspher_tangential_create": (i, [i, d, d, d, C.POINTER(vp)]),
        "aspher_tangential_destroy": (None, [vp]),
        "aspher_tangential_matvec": (i, [vp, dp, dp]),
        "aspher_tangential_matvec_device": (i, [vp, vp, vp, i]),
        "aspher_tangential_precond_apply": (i, [i, d, i, dp, C.POINTER(C.c_uint8), i, dp]),
        "aspher_cerruti_points": (i, [i, dp, dp, dp, dp, d, dp, dp]),
        "aspher_cerruti_tables": (i, [i, d, d, d, dp, dp]),
        "aspher_rough_surface_device": (i, [i, d, d, d, d, C.c_uint64, vp]),
        "aspher_solve_ensemble": (i, [i, ip, ip, dp, dp, dp, C.POINTER(dp),
                                      C.POINTER(EnsembleParams), C.POINTER(vp),
                                      C.POINTER(EnsembleStats), dp]),
        "aspher_solve_ensemble_device": (i, [i, ip, ip, dp, dp, dp, C.POINTER(vp),
                                             C.POINTER(EnsembleParams), C.POINTER(vp),
                                             C.POINTER(EnsembleStats), dp]),
        "aspher_adjoint_params_default": (None, [C.POINTER(AdjointParams)]),
        "aspher_solver_adjoint": (i, [vp, vp, dp, C.POINTER(AdjointParams), dp,
                                      C.POINTER(AdjointInfo)]),
        "aspher_solver_adjoint_device": (i, [vp, vp, vp, C.POINTER(AdjointParams), vp,
                                             C.POINTER(AdjointInfo)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def lib():
    """Load libaspher_hip.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libaspher_hip.so is not built (%s); run `python -m aspher_amd.build`"
                % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def check(rc: int) -> None:
    """Map a C-ABI status code to the reference's Python exception classes."""
    if rc == OK:
        return
    msg = lib().aspher_last_error().decode(errors="replace")
    if rc == ERR_INVALID:
        raise ValueError(msg)
    raise RuntimeError(msg)


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))
