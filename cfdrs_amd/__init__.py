"""cfdrs_amd — MI355X-native sparse linear-solve path of CFDrs.

GPU library ``libcfdla.so`` (hand-written gfx950 HIP kernels behind a C ABI,
``include/cfdla.h``) with a Python mirror of the reference's Krylov and
preconditioner interfaces.
"""
from . import _native
from ._native import CfdlaError, NativeLibraryMissing
from .linear_solver import (IterativeSolverConfig, KrylovWorkspace, Precision, SolveError, SolveOutcome,
                            SolveReport, SolverError, SolverKind, Termination, attempt, bicgstab,
                            bicgstab_preconditioned, cg, cg_preconditioned, converged_or_none, gmres,
                            gmres_preconditioned, gmres_preconditioned_with_workspace, interpret, restart_width,
                            solve_transpose)
from .multigrid import (AlgebraicMultigrid, AMGConfig, AMGStatistics, CoarseningStrategy, CoarseSolver,
                        InterpolationStrategy, JacobiPreconditioner, SmootherType, SuccessiveOverRelaxation)
from .sparse import CsrMatrix, SparseMatrixBuilder
from . import simple

__all__ = [
    "AlgebraicMultigrid", "AMGConfig", "AMGStatistics", "CfdlaError", "CoarseningStrategy", "CoarseSolver",
    "CsrMatrix", "InterpolationStrategy", "IterativeSolverConfig", "JacobiPreconditioner", "KrylovWorkspace",
    "NativeLibraryMissing", "Precision", "SmootherType", "SolveError", "SolveOutcome", "SolveReport",
    "SolverError", "SolverKind", "SparseMatrixBuilder", "SuccessiveOverRelaxation", "Termination", "attempt",
    "bicgstab",
    "bicgstab_preconditioned", "cg", "cg_preconditioned", "converged_or_none", "gmres", "gmres_preconditioned",
    "gmres_preconditioned_with_workspace", "interpret", "restart_width", "solve_transpose",
]
