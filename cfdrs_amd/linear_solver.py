"""Krylov entry points over the GPU library.

Mirrors ``cfd_math::linear_solver::krylov``
(crates/cfd-math/src/linear_solver/krylov.rs): the solve functions
(:150-373), the restart ladder (:84-114), ``KrylovWorkspace`` (:35-81),
``interpret`` / ``SolveOutcome`` / ``converged_or_none`` (:375-523) and
``SolverKind`` (:534-613).  Solutions are numpy arrays (copied to and from
the device) or torch CUDA float64 tensors (used in place).

Numerical termination is value-semantic, as in the reference: a stalled or
broken-down solve returns a ``SolveReport``; only dimension, configuration
and device failures raise ``SolveError``.
"""
from __future__ import annotations

import ctypes as C
import enum
import logging
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .sparse import CsrMatrix, _is_torch, bind_torch_stream, require_device_vectors

log = logging.getLogger("cfdrs_amd")


class Termination(enum.IntEnum):
    Converged = N.TERM_CONVERGED
    InitialResidual = N.TERM_INITIAL_RESIDUAL
    MaxIterations = N.TERM_MAX_ITERATIONS
    Breakdown = N.TERM_BREAKDOWN
    NonPositiveCurvature = N.TERM_NONPOSITIVE_CURVATURE


class Precision(enum.IntEnum):
    FP64 = N.FP64
    FP32_REFINED = N.FP32_REFINED
    # FP32-refined when an iteration streams more than 256 MiB of FP64
    # operator bytes (matrix and preconditioner), FP64 otherwise; batched
    # network solves take FP64.  SolveReport.precision says which ran.
    AUTO = N.PRECISION_AUTO


@dataclass
class SolveReport:
    """athena SolveReport as read at krylov.rs:506-516."""

    iterations: int
    initial_residual_norm: float
    final_residual_norm: float
    threshold: float
    termination: Termination
    outer_iterations: int = 0
    precision: Precision = Precision.FP64  # the arithmetic the solve ran in


class SolveError(RuntimeError):
    """Dimension, configuration or backend failure (SolveError<LetoBackendError>)."""

    def __init__(self, err: N.CfdlaError):
        super().__init__(str(err))
        self.status = err.status


class SolverError(RuntimeError):
    """cfd_core::error::Error::Solver raised by ``interpret``."""


@dataclass
class IterativeSolverConfig:
    """Stopping rule ||b - A x|| <= max(absolute_tolerance, tolerance * ||b||).

    ``IterativeSolverConfig::new(tol)`` maps to a relative tolerance; the
    reference's ext leto-ops semantics are not available to pin.
    """

    tolerance: float = 1e-6
    absolute_tolerance: float = 0.0
    max_iterations: int = 1000
    use_preconditioner: bool = False
    precision: Precision = Precision.FP64
    inner_tolerance: float = 1e-5
    max_outer: int = 30
    check_every: int = 0
    reliable_interval: int = 0
    reliable_delta: float = 0.0

    @classmethod
    def new(cls, tolerance: float) -> "IterativeSolverConfig":
        return cls(tolerance=tolerance)

    def with_max_iterations(self, n: int) -> "IterativeSolverConfig":
        self.max_iterations = int(n)
        return self

    def with_preconditioner(self) -> "IterativeSolverConfig":
        self.use_preconditioner = True
        return self

    def with_precision(self, p: Precision, inner_tolerance: float | None = None) -> "IterativeSolverConfig":
        self.precision = Precision(p)
        if inner_tolerance is not None:
            self.inner_tolerance = inner_tolerance
        return self

    def with_reliable_interval(self, k: int) -> "IterativeSolverConfig":
        """FP32_REFINED CG: FP64 residual replacement every k iterations
        (0 = when the residual has dropped by ``reliable_delta``); -1 selects
        restarted refinement."""
        self.reliable_interval = int(k)
        return self

    def with_reliable_delta(self, delta: float) -> "IterativeSolverConfig":
        """FP32_REFINED CG: residual reduction between FP64 residual
        replacements (0 = library default, 1e-1)."""
        self.reliable_delta = float(delta)
        return self

    def native(self, kind: int, restart: int = 30, ladder: bool = True) -> N.SolverConfig:
        return N.SolverConfig(kind, restart, 1 if ladder else 0, int(self.precision), self.absolute_tolerance,
                              self.tolerance, self.max_iterations, self.inner_tolerance, self.max_outer,
                              self.check_every, self.reliable_interval, 0, self.reliable_delta)


def restart_width(requested: int) -> int:
    """Smallest ladder width covering ``requested`` (krylov.rs:99-113)."""
    for w in (8, 16, 32, 64, 128):
        if requested <= w:
            return w
    return 256


class KrylovWorkspace:
    """Reusable GMRES workspace descriptor (krylov.rs:35-81): the restart is
    rounded up the ladder; device vectors are pooled by the library."""

    def __init__(self, restart: int, dimension: int):
        self.width = restart_width(restart)
        self.dimension = dimension


def _solve(matrix: CsrMatrix, rhs, preconditioner, solution, config: IterativeSolverConfig, kind: int,
           restart: int = 30, ladder: bool = True, transpose: bool = False) -> SolveReport:
    cfg = config.native(kind, restart, ladder)
    rep = N.Report()
    prec = preconditioner.handle if preconditioner is not None else None
    n = matrix.nrows
    try:
        if _is_torch(solution):
            import torch

            require_device_vectors(rhs, solution)
            if solution.dtype != torch.float64 or rhs.dtype != torch.float64:
                raise N.CfdlaError(N.ERR_INVALID_ARG, "device solves take float64 tensors")
            if solution.numel() != n or rhs.numel() != n:
                raise N.CfdlaError(N.ERR_DIMENSION, "vector length differs from matrix dimension")
            bind_torch_stream()
            fn = "cfdla_solve_transpose_device" if transpose else "cfdla_solve_device"
            N.call(fn, matrix.handle, prec, rhs.data_ptr(), solution.data_ptr(), C.byref(cfg), C.byref(rep))
        else:
            if transpose:
                raise N.CfdlaError(N.ERR_INVALID_ARG, "transpose solves take device tensors")
            b = np.ascontiguousarray(rhs, dtype=np.float64)
            if b.shape != (n,) or solution.shape != (n,):
                raise N.CfdlaError(N.ERR_DIMENSION, "vector length differs from matrix dimension")
            if solution.dtype != np.float64 or not solution.flags.c_contiguous:
                raise N.CfdlaError(N.ERR_INVALID_ARG, "solution must be a contiguous float64 array")
            N.call("cfdla_solve", matrix.handle, prec, N.ptr(b), N.ptr(solution), C.byref(cfg), C.byref(rep))
    except N.CfdlaError as e:
        raise SolveError(e) from None
    return SolveReport(rep.iterations, rep.initial_residual_norm, rep.final_residual_norm, rep.threshold,
                       Termination(rep.termination), rep.outer_iterations, Precision(rep.precision))


def cg_preconditioned(matrix, right_hand_side, preconditioner, solution, config) -> SolveReport:
    """krylov::cg_preconditioned (krylov.rs:318-343)."""
    return _solve(matrix, right_hand_side, preconditioner, solution, config, N.KIND_CG)


def cg(matrix, right_hand_side, solution, config) -> SolveReport:
    """krylov::cg (krylov.rs:345-360)."""
    return _solve(matrix, right_hand_side, None, solution, config, N.KIND_CG)


def bicgstab_preconditioned(matrix, right_hand_side, preconditioner, solution, config) -> SolveReport:
    """krylov::bicgstab_preconditioned (krylov.rs:253-278)."""
    return _solve(matrix, right_hand_side, preconditioner, solution, config, N.KIND_BICGSTAB)


def bicgstab(matrix, right_hand_side, solution, config) -> SolveReport:
    """krylov::bicgstab (krylov.rs:280-316)."""
    return _solve(matrix, right_hand_side, None, solution, config, N.KIND_BICGSTAB)


def gmres_preconditioned(matrix, right_hand_side, preconditioner, solution, config, restart: int) -> SolveReport:
    """krylov::gmres_preconditioned (krylov.rs:150-173); restart rounded up the ladder."""
    return _solve(matrix, right_hand_side, preconditioner, solution, config, N.KIND_GMRES, restart, True)


def gmres_preconditioned_with_workspace(matrix, right_hand_side, preconditioner, solution, config,
                                        workspace: KrylovWorkspace) -> SolveReport:
    """krylov::gmres_preconditioned_with_workspace (krylov.rs:175-229)."""
    if workspace.dimension != matrix.nrows:
        raise SolveError(N.CfdlaError(N.ERR_DIMENSION, "workspace dimension differs from matrix"))
    return _solve(matrix, right_hand_side, preconditioner, solution, config, N.KIND_GMRES, workspace.width, False)


def gmres(matrix, right_hand_side, solution, config, restart: int) -> SolveReport:
    """krylov::gmres (krylov.rs:231-251)."""
    return _solve(matrix, right_hand_side, None, solution, config, N.KIND_GMRES, restart, True)


def gmres_exact_restart(matrix, right_hand_side, preconditioner, solution, config, restart: int) -> SolveReport:
    """GMRES(m) with exactly the requested m (no ladder rounding)."""
    return _solve(matrix, right_hand_side, preconditioner, solution, config, N.KIND_GMRES, restart, False)


def solve_transpose(matrix, right_hand_side, preconditioner, solution, config,
                    kind: int = N.KIND_CG) -> SolveReport:
    """A^T lambda = g on device tensors (adjoint solves)."""
    return _solve(matrix, right_hand_side, preconditioner, solution, config, kind, transpose=True)


class SolveOutcome:
    """krylov::SolveOutcome (krylov.rs:404-431)."""

    Converged, Stalled, BrokenDown = "Converged", "Stalled", "BrokenDown"

    def __init__(self, kind: str, report: SolveReport):
        self.kind, self._report = kind, report

    def report(self) -> SolveReport:
        return self._report

    def converged(self) -> bool:
        return self.kind == SolveOutcome.Converged

    def __repr__(self):
        return f"SolveOutcome.{self.kind}({self._report})"


def interpret(context: str, outcome) -> SolveOutcome:
    """krylov::interpret (krylov.rs:375-400).  ``outcome`` is a SolveReport or
    the SolveError a solve raised."""
    if isinstance(outcome, BaseException):
        raise SolverError(f"{context}: linear solve failed: {outcome}")
    t = outcome.termination
    if t in (Termination.Converged, Termination.InitialResidual):
        return SolveOutcome(SolveOutcome.Converged, outcome)
    if t == Termination.MaxIterations:
        return SolveOutcome(SolveOutcome.Stalled, outcome)
    if t in (Termination.Breakdown, Termination.NonPositiveCurvature):
        return SolveOutcome(SolveOutcome.BrokenDown, outcome)
    raise SolverError(f"{context}: linear solve terminated as {t!r}")


def converged_or_none(context: str, outcome) -> SolveReport | None:
    """krylov::converged_or_none (krylov.rs:497-523)."""
    try:
        o = interpret(context, outcome)
    except SolverError as e:
        log.warning("%s", e)
        return None
    if o.converged():
        return o.report()
    r = o.report()
    log.warning("%s: did not converge (%s) after %d iterations, residual %g > %g", context, r.termination.name,
                r.iterations, r.final_residual_norm, r.threshold)
    return None


def attempt(fn, *args):
    """Run a solve and return its SolveReport, or the SolveError it raised
    (the Python spelling of KrylovResult for ``interpret``)."""
    try:
        return fn(*args)
    except SolveError as e:
        return e


class SolverKind:
    """krylov::SolverKind (krylov.rs:534-613)."""

    def __init__(self, tag: str, restart: int = 30):
        self.tag, self.restart = tag, restart

    @classmethod
    def ConjugateGradient(cls):
        return cls("ConjugateGradient")

    @classmethod
    def BiCgStab(cls):
        return cls("BiCGSTAB")

    @classmethod
    def Gmres(cls, restart: int):
        return cls("GMRES", restart)

    def name(self) -> str:
        return self.tag

    def solve(self, matrix, right_hand_side, solution, config) -> SolveReport:
        if self.tag == "ConjugateGradient":
            return cg(matrix, right_hand_side, solution, config)
        if self.tag == "BiCGSTAB":
            return bicgstab(matrix, right_hand_side, solution, config)
        return gmres(matrix, right_hand_side, solution, config, self.restart)

    def solve_preconditioned(self, matrix, right_hand_side, preconditioner, solution, config) -> SolveReport:
        if self.tag == "ConjugateGradient":
            return cg_preconditioned(matrix, right_hand_side, preconditioner, solution, config)
        if self.tag == "BiCGSTAB":
            return bicgstab_preconditioned(matrix, right_hand_side, preconditioner, solution, config)
        return gmres_preconditioned(matrix, right_hand_side, preconditioner, solution, config, self.restart)

    def __eq__(self, other):
        return isinstance(other, SolverKind) and (self.tag, self.restart if self.tag == "GMRES" else 0) == (
            other.tag, other.restart if other.tag == "GMRES" else 0)

    def __repr__(self):
        return f"SolverKind.{self.tag}" + (f"(restart={self.restart})" if self.tag == "GMRES" else "")
