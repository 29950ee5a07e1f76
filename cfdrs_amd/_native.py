"""ctypes bindings of libcfdla.so (the C ABI declared in include/cfdla.h).

The shared library is built in-tree (``cfdrs_amd/libcfdla.so``) by
``__graft_entry__.build()`` or ``make -C cfdrs_amd/csrc``.  There is no
fallback: if the library is missing, importing the solver entry points
raises ``NativeLibraryMissing``.
"""
from __future__ import annotations

import ctypes as C
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcfdla.so")

# status codes (cfdla_status)
OK, ERR_INVALID_ARG, ERR_DIMENSION, ERR_CONFIG, ERR_NUMERICAL, ERR_DEVICE, ERR_OOM = range(7)
# termination (cfdla_termination)
TERM_CONVERGED, TERM_INITIAL_RESIDUAL, TERM_MAX_ITERATIONS, TERM_BREAKDOWN, TERM_NONPOSITIVE_CURVATURE = range(5)
# solver kinds / precision / dtype
KIND_CG, KIND_BICGSTAB, KIND_GMRES = range(3)
FP64, FP32_REFINED, PRECISION_AUTO = range(3)
DT_F64, DT_F32 = range(2)
MAX_LEVELS = 32


class NativeLibraryMissing(RuntimeError):
    """libcfdla.so is not built; run __graft_entry__.build()."""


class CfdlaError(RuntimeError):
    """A libcfdla call failed; ``status`` holds the cfdla_status code."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[cfdla status {status}] {message}")
        self.status = status


class SolverConfig(C.Structure):
    _fields_ = [("kind", C.c_int32), ("restart", C.c_int32), ("restart_ladder", C.c_int32),
                ("precision", C.c_int32), ("tol_abs", C.c_double), ("tol_rel", C.c_double),
                ("max_iter", C.c_int64), ("inner_tol", C.c_double), ("max_outer", C.c_int32),
                ("check_every", C.c_int32), ("reliable_interval", C.c_int32), ("reserved", C.c_int32),
                ("reliable_delta", C.c_double)]


class Report(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("outer_iterations", C.c_int64),
                ("initial_residual_norm", C.c_double), ("final_residual_norm", C.c_double),
                ("threshold", C.c_double), ("termination", C.c_int32), ("precision", C.c_int32)]


class AmgConfig(C.Structure):
    _fields_ = [("max_levels", C.c_int32), ("min_coarse_size", C.c_int32), ("coarsening", C.c_int32),
                ("interpolation", C.c_int32), ("smoother", C.c_int32), ("pre_smooth_iterations", C.c_int32),
                ("post_smooth_iterations", C.c_int32), ("coarse_solver", C.c_int32),
                ("relaxation_factor", C.c_double), ("strength_threshold", C.c_double),
                ("smoothing_omega", C.c_double), ("max_aggregate_size", C.c_int32),
                ("max_interpolation_points", C.c_int32), ("galerkin_on_device", C.c_int32),
                ("separate_coarse_restriction", C.c_int32)]


class AmgStatistics(C.Structure):
    _fields_ = [("num_levels", C.c_int32), ("fused_level", C.c_int32),
                ("level_rows", C.c_int64 * MAX_LEVELS), ("level_nnz", C.c_int64 * MAX_LEVELS),
                ("operator_complexity", C.c_double), ("grid_complexity", C.c_double),
                ("setup_time", C.c_double)]


class BatchSchedule(C.Structure):
    _fields_ = [("wave_max_rows", C.c_int32), ("blocks_per_cu", C.c_int32)]


_VP = C.c_void_p
_I64P = C.POINTER(C.c_int64)
_F64P = C.POINTER(C.c_double)
_I32P = C.POINTER(C.c_int32)

# name -> (restype, argtypes); every function returns cfdla_status unless noted
SIGNATURES = {
    "cfdla_last_error": (C.c_char_p, []),
    "cfdla_version": (C.c_char_p, []),
    "cfdla_init": (C.c_int, [C.c_int32]),
    "cfdla_set_stream": (C.c_int, [_VP]),
    "cfdla_synchronize": (C.c_int, []),
    "cfdla_matrix_create": (C.c_int, [C.c_int64, C.c_int64, _VP, _VP, _VP, C.POINTER(_VP)]),
    "cfdla_matrix_create_device": (C.c_int, [C.c_int64, C.c_int64, C.c_int64, _VP, _VP, _VP, C.POINTER(_VP)]),
    "cfdla_matrix_update_values": (C.c_int, [_VP, _VP]),
    "cfdla_matrix_update_values_device": (C.c_int, [_VP, _VP]),
    "cfdla_matrix_get_values": (C.c_int, [_VP, _VP]),
    "cfdla_matrix_info": (C.c_int, [_VP, _I64P, _I64P, _I64P]),
    "cfdla_matrix_storage": (C.c_int, [_VP, _I32P, _I64P, _I64P]),
    "cfdla_matrix_set_storage": (C.c_int, [_VP, C.c_int32]),
    "cfdla_matrix_set_gather": (C.c_int, [_VP, C.c_int32]),
    "cfdla_matrix_gather": (C.c_int, [_VP, _I32P, _I32P]),
    "cfdla_matrix_destroy": (C.c_int, [_VP]),
    "cfdla_spmv": (C.c_int, [_VP, _VP, _VP]),
    "cfdla_spmv_device": (C.c_int, [_VP, C.c_int32, _VP, _VP]),
    "cfdla_residual_device": (C.c_int, [_VP, _VP, _VP, _VP, _F64P]),
    "cfdla_precond_create_jacobi": (C.c_int, [_VP, C.POINTER(_VP)]),
    "cfdla_precond_create_amg": (C.c_int, [_VP, C.POINTER(AmgConfig), C.POINTER(_VP)]),
    "cfdla_precond_create_sor": (C.c_int, [_VP, C.c_double, _I32P, C.POINTER(_VP)]),
    "cfdla_precond_apply": (C.c_int, [_VP, _VP, _VP]),
    "cfdla_precond_apply_device": (C.c_int, [_VP, C.c_int32, _VP, _VP]),
    "cfdla_amg_update_values": (C.c_int, [_VP, _VP]),
    "cfdla_amg_set_gather": (C.c_int, [_VP, C.c_int32]),
    "cfdla_amg_update_values_device": (C.c_int, [_VP, _VP]),
    "cfdla_amg_stats": (C.c_int, [_VP, C.POINTER(AmgStatistics)]),
    "cfdla_amg_export_level": (C.c_int, [_VP, C.c_int32, C.c_int32, _I64P, _I64P, _I64P, _VP, _VP, _VP]),
    "cfdla_amg_export_map": (C.c_int, [_VP, C.c_int32, _VP]),
    "cfdla_precond_destroy": (C.c_int, [_VP]),
    "cfdla_solve": (C.c_int, [_VP, _VP, _VP, _VP, C.POINTER(SolverConfig), C.POINTER(Report)]),
    "cfdla_solve_device": (C.c_int, [_VP, _VP, _VP, _VP, C.POINTER(SolverConfig), C.POINTER(Report)]),
    "cfdla_solve_transpose_device": (C.c_int, [_VP, _VP, _VP, _VP, C.POINTER(SolverConfig), C.POINTER(Report)]),
    "cfdla_coef_grad_poisson2d_device": (C.c_int, [C.c_int64, C.c_int64, C.c_double, C.c_int32, _VP, _VP, _VP,
                                                   _VP, _VP, _VP]),
    "cfdla_synth_uniform": (C.c_int, [C.c_uint64, C.c_double, C.c_double, C.c_int64, _VP]),
    "cfdla_synth_grf": (C.c_int, [C.c_int64, C.c_int64, C.c_double, C.c_uint64, _VP]),
    "cfdla_problem_momentum_u": (C.c_int, [C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double,
                                           C.c_double, _VP, _VP, _I64P, _VP, _VP, _VP]),
    "cfdla_problem_pressure_rhs": (C.c_int, [C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double,
                                             _VP, _VP, C.c_int32, _VP]),
    "cfdla_problem_simple_pressure": (C.c_int, [C.c_int64, C.c_int64, C.c_double, C.c_double, _VP, C.c_int32, _VP,
                                                _VP, _VP, _VP, _VP, _VP, _I64P, _VP, _VP, _VP, _VP, _F64P]),
    "cfdla_simple_pressure_refresh_device": (C.c_int, [_VP, C.c_int64, C.c_int64, C.c_double, C.c_double,
                                                       C.c_int32, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _F64P]),
    "cfdla_simple_corrections_device": (C.c_int, [C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double, _VP,
                                                  _VP, _VP, _VP, _VP, _VP]),
    "cfdla_pressure_rhs_device": (C.c_int, [C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double,
                                            _VP, _VP, C.c_int32, _VP]),
    "cfdla_problem_pressure_correction": (C.c_int, [C.c_int64, C.c_int64, C.c_double, C.c_double, _VP, _I64P,
                                                    _VP, _VP, _VP]),
    "cfdla_problem_varcoef": (C.c_int, [C.c_int64, C.c_int64, C.c_int64, C.c_double, _VP, C.c_int32, _VP, _I64P,
                                        _VP, _VP, _VP, _VP]),
    "cfdla_batch_create": (C.c_int, [C.c_int64, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, C.POINTER(_VP)]),
    "cfdla_batch_update_conductance": (C.c_int, [_VP, _VP]),
    "cfdla_batch_update_conductance_device": (C.c_int, [_VP, _VP]),
    "cfdla_batch_info": (C.c_int, [_VP, _I64P, _I64P, _I64P]),
    "cfdla_batch_export": (C.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "cfdla_batch_set_schedule": (C.c_int, [_VP, C.POINTER(BatchSchedule)]),
    "cfdla_batch_get_schedule": (C.c_int, [_VP, C.POINTER(BatchSchedule)]),
    "cfdla_batch_solve": (C.c_int, [_VP, C.POINTER(SolverConfig), _VP, _VP]),
    "cfdla_batch_solve_device": (C.c_int, [_VP, C.POINTER(SolverConfig), _VP, _VP, _VP]),
    "cfdla_batch_edge_flows_device": (C.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "cfdla_batch_outlet_flux_device": (C.c_int, [_VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "cfdla_varcoef_refresh_device": (C.c_int, [_VP, C.c_int64, C.c_int64, C.c_int64, C.c_double, _VP, C.c_int32,
                                               _VP, _VP]),
    "cfdla_batch_solve_device_noalloc": (C.c_int, [_VP, C.POINTER(SolverConfig), _VP, _VP, _VP]),
    "cfdla_batch_autotune": (C.c_int, [_VP, C.POINTER(SolverConfig), C.c_int32, _VP, _VP, C.POINTER(BatchSchedule)]),
    "cfdla_batch_edge_gradient_device": (C.c_int, [_VP, _VP, _VP, _VP, _VP, _VP]),
    "cfdla_batch_destroy": (C.c_int, [_VP]),
    "cfdla_synth_networks": (C.c_int, [C.c_int64, C.c_uint64, C.c_int64, C.c_int64, _VP, _VP, _VP, _VP, _VP, _VP]),
    "cfdla_rect_conductance": (C.c_int, [C.c_int64, _VP, _VP, _VP, C.c_double, _VP, _VP]),
    "cfdla_rect_conductance_device": (C.c_int, [C.c_int64, _VP, _VP, _VP, C.c_double, _VP, _VP]),
}

_lib = None


def lib():
    """Load libcfdla.so (no fallback: raises NativeLibraryMissing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() or make -C cfdrs_amd/csrc")
        # One HIP runtime per process: torch ships its own libamdhip64 /
        # libhsa-runtime64.  Loaded first, they satisfy libcfdla's
        # libamdhip64.so.7 dependency by soname, so both share one runtime
        # (and torch's stream handles are valid here).  Loaded after
        # libcfdla, torch would bring a second HSA runtime that finds no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != OK:
        msg = lib().cfdla_last_error().decode(errors="replace")
        raise CfdlaError(status, msg)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))


def ptr(a) -> int | None:
    """Raw address of a numpy array or torch tensor (None passes NULL)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data
