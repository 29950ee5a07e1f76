"""Preconditioners: AMG and Jacobi on the GPU.

Mirrors ``cfd_math::multigrid``:
``AMGConfig`` and its defaults (crates/cfd-math/src/linear_solver/preconditioners/multigrid/mod.rs:219-260),
``AlgebraicMultigrid::{new, with_config, recompute, apply_to}`` (amg.rs:133-558),
``AMGStatistics`` (mod.rs:366-395), and ``JacobiPreconditioner``
(leto-ops, re-exported at crates/cfd-math/src/lib.rs:181-184).

The hierarchy is set up on the host (bit-exact integer structure) and the
V-cycle runs on the device.  The device V-cycle smooths with Jacobi only:
Gauss-Seidel / SGS / SOR are sequential sweeps and are rejected with a
configuration error rather than silently replaced.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .sparse import CsrMatrix, _is_torch, bind_torch_stream, require_device_vectors


class CoarseningStrategy(enum.IntEnum):
    RugeStueben = 0
    Aggregation = 1
    Hybrid = 2
    # three-phase aggregation (Vanek, Mandel & Brezina 1996); no reference
    # counterpart: the reference's greedy aggregation stalls on strongly
    # varying coefficients (coarsening ratio -> 1 after a few levels)
    ThreePhaseAggregation = 3


class InterpolationStrategy(enum.IntEnum):
    Classical = 0
    Direct = 1
    SmoothedDirect = 2  # P = (I - omega D^-1 A) P_direct; no reference counterpart


class SmootherType(enum.IntEnum):
    GaussSeidel = 0
    SymmetricGaussSeidel = 1
    Jacobi = 2
    SOR = 3
    Chebyshev = 4


class CoarseSolver(enum.IntEnum):
    ReferenceSweeps = 0  # pre-smoothing + 10 sweeps (amg.rs:501)
    Direct = 1           # FP64 dense direct solve


class CycleType(enum.Enum):
    VCycle = "VCycle"


@dataclass
class AMGConfig:
    max_levels: int = 10
    min_coarse_size: int = 50
    coarsening_strategy: CoarseningStrategy = CoarseningStrategy.RugeStueben
    interpolation_strategy: InterpolationStrategy = InterpolationStrategy.Classical
    cycle_type: CycleType = CycleType.VCycle
    smoother_type: SmootherType = SmootherType.GaussSeidel
    pre_smooth_iterations: int = 2
    post_smooth_iterations: int = 2
    relaxation_factor: float = 1.0
    strength_threshold: float = 0.25
    max_interpolation_points: int = 4
    coarse_solver: CoarseSolver = CoarseSolver.ReferenceSweeps
    smoothing_omega: float = 2.0 / 3.0
    max_aggregate_size: int = 8
    # form the Galerkin products on the GPU (bit-identical to the host
    # products; no reference counterpart)
    galerkin_on_device: bool = False
    # keep the pre-sweeps, residual and restriction of the level above a
    # dense coarsest level as three products instead of one fused product
    # (the same operator in exact arithmetic; no reference counterpart)
    separate_coarse_restriction: bool = False

    @classmethod
    def default(cls) -> "AMGConfig":
        return cls()

    @classmethod
    def gpu_aggregation(cls, smoothed: bool = True, relaxation_factor: float = 0.8) -> "AMGConfig":
        """The reference's greedy aggregation (theta 0.5, size <= 8), Jacobi 2/2,
        FP64 direct coarsest solve."""
        return cls(coarsening_strategy=CoarseningStrategy.Aggregation,
                   interpolation_strategy=(InterpolationStrategy.SmoothedDirect if smoothed
                                           else InterpolationStrategy.Direct),
                   smoother_type=SmootherType.Jacobi, relaxation_factor=relaxation_factor,
                   coarse_solver=CoarseSolver.Direct)

    @classmethod
    def smoothed_aggregation(cls, strength_threshold: float = 0.25, relaxation_factor: float = 0.8,
                             smoothed: bool = True) -> "AMGConfig":
        """Three-phase aggregates, smoothed prolongator, Jacobi 2/2, FP64
        direct coarsest solve once a level has at most 1000 rows (the 128^3
        benchmark recipe: 2097152 -> 271768 -> 16468 -> 641 rows)."""
        return cls(coarsening_strategy=CoarseningStrategy.ThreePhaseAggregation, min_coarse_size=1000,
                   interpolation_strategy=(InterpolationStrategy.SmoothedDirect if smoothed
                                           else InterpolationStrategy.Direct),
                   smoother_type=SmootherType.Jacobi, relaxation_factor=relaxation_factor,
                   strength_threshold=strength_threshold, coarse_solver=CoarseSolver.Direct)

    def native(self) -> N.AmgConfig:
        if self.cycle_type is not CycleType.VCycle:
            raise N.CfdlaError(N.ERR_CONFIG, "only the V-cycle is implemented")
        return N.AmgConfig(self.max_levels, self.min_coarse_size, int(self.coarsening_strategy),
                           int(self.interpolation_strategy), int(self.smoother_type), self.pre_smooth_iterations,
                           self.post_smooth_iterations, int(self.coarse_solver), self.relaxation_factor,
                           self.strength_threshold, self.smoothing_omega, self.max_aggregate_size,
                           self.max_interpolation_points, 1 if self.galerkin_on_device else 0,
                           1 if self.separate_coarse_restriction else 0)


@dataclass
class AMGStatistics:
    num_levels: int
    level_sizes: list = field(default_factory=list)
    level_nnz: list = field(default_factory=list)
    operator_complexity: float = 1.0
    grid_complexity: float = 1.0
    setup_time: float = 0.0
    # the level whose pre-sweeps and restriction run as one product (-1:
    # none, or the hierarchy is not on the device yet)
    fused_level: int = -1


class _Preconditioner:
    _h = None

    @property
    def handle(self):
        if self._h is None:
            raise ValueError("preconditioner has been closed")
        return self._h

    def apply_to(self, r, z) -> None:
        """z = M^-1 r (z fully overwritten); numpy or torch FP64/FP32 vectors."""
        if _is_torch(r):
            import torch

            require_device_vectors(r, z)
            dt = {torch.float64: N.DT_F64, torch.float32: N.DT_F32}.get(r.dtype)
            if dt is None or z.dtype != r.dtype:
                raise N.CfdlaError(N.ERR_INVALID_ARG, "r and z must be float64 or float32 tensors of one dtype")
            if r.numel() != self.n or z.numel() != self.n:
                raise N.CfdlaError(N.ERR_DIMENSION, "AMG preconditioner vector length mismatch")
            bind_torch_stream()
            N.call("cfdla_precond_apply_device", self.handle, dt, r.data_ptr(), z.data_ptr())
            return
        rr = np.ascontiguousarray(r, dtype=np.float64)
        if rr.shape != (self.n,) or z.shape != (self.n,):
            raise N.CfdlaError(N.ERR_DIMENSION, f"preconditioner vector length mismatch: {rr.shape} vs {self.n}")
        out = np.empty(self.n)
        N.call("cfdla_precond_apply", self.handle, N.ptr(rr), N.ptr(out))
        z[:] = out

    def close(self) -> None:
        if self._h is not None and N._lib is not None:
            N.lib().cfdla_precond_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class JacobiPreconditioner(_Preconditioner):
    """z = D^-1 r; identity where |d_ii| <= f64::EPSILON
    (crates/cfd-1d/src/solver/core/linear_system.rs:270-299)."""

    def __init__(self, matrix: CsrMatrix):
        h = C.c_void_p()
        N.call("cfdla_precond_create_jacobi", matrix.handle, C.byref(h))
        self._h, self.n, self.matrix = h, matrix.nrows, matrix


class SuccessiveOverRelaxation(_Preconditioner):
    """SOR preconditioner z = (D / omega + L)^-1 r of the momentum solve
    (momentum/solve.rs:137-147, ext athena SuccessiveOverRelaxation), swept in
    multicolour order on the device: rows coloured greedily in breadth-first
    order (red-black for 5-point stencils), one parallel forward sweep per
    colour; rows without a diagonal take a unit pivot, as
    from_csr_with_identity_rows does."""

    def __init__(self, matrix: CsrMatrix, omega: float = 1.0):
        h = C.c_void_p()
        nc = C.c_int32()
        N.call("cfdla_precond_create_sor", matrix.handle, float(omega), C.byref(nc), C.byref(h))
        self._h, self.n, self.matrix, self.omega, self.colours = h, matrix.nrows, matrix, float(omega), nc.value

    @classmethod
    def from_csr_with_identity_rows(cls, matrix: CsrMatrix, omega: float) -> "SuccessiveOverRelaxation":
        return cls(matrix, omega)


class AlgebraicMultigrid(_Preconditioner):
    """AMG preconditioner (amg.rs:100-558) with a device V-cycle."""

    def __init__(self, matrix: CsrMatrix, config: AMGConfig | None = None):
        self.config = config or AMGConfig.default()
        cfg = self.config.native()
        h = C.c_void_p()
        N.call("cfdla_precond_create_amg", matrix.handle, C.byref(cfg), C.byref(h))
        self._h, self.n, self.matrix = h, matrix.nrows, matrix

    @classmethod
    def new(cls, matrix, config):
        return cls(matrix, config)

    with_config = new

    def recompute(self, fine_matrix: CsrMatrix) -> None:
        """AlgebraicMultigrid::recompute (amg.rs:179-214)."""
        if fine_matrix.nnz != self.matrix.nnz:
            raise N.CfdlaError(N.ERR_DIMENSION, "recompute requires the same sparsity pattern")
        # keep the array referenced across the call: for a matrix created with
        # unsorted rows creation_order_values() is a fresh array
        values = np.ascontiguousarray(fine_matrix.creation_order_values(), dtype=np.float64)
        N.call("cfdla_amg_update_values", self.handle, N.ptr(values))

    def recompute_device(self, fine_matrix: CsrMatrix) -> None:
        """AlgebraicMultigrid::recompute (amg.rs:179-214) from the current
        device values of the matrix the hierarchy was built from, on the GPU
        (bit-identical to ``recompute`` with the same values)."""
        if fine_matrix is not self.matrix:
            raise N.CfdlaError(N.ERR_INVALID_ARG, "recompute_device needs the matrix the hierarchy was built from")
        N.call("cfdla_amg_update_values_device", self.handle, fine_matrix.handle)

    def set_gather(self, layout: str) -> None:
        """Gather layout of every level's row-block CSR products: "auto"
        (chosen per operator), "quad", "rows", "strided" or "sorted" (see
        CsrMatrix.set_gather)."""
        from .sparse import CsrMatrix as _Csr

        if layout not in _Csr.GATHER_LAYOUTS:
            raise ValueError(f"unknown gather layout {layout!r}")
        N.call("cfdla_amg_set_gather", self.handle, _Csr.GATHER_LAYOUTS[layout])

    @property
    def statistics(self) -> AMGStatistics:
        s = N.AmgStatistics()
        N.call("cfdla_amg_stats", self.handle, C.byref(s))
        L = s.num_levels
        return AMGStatistics(L, list(s.level_rows[:L]), list(s.level_nnz[:L]), s.operator_complexity,
                             s.grid_complexity, s.setup_time, s.fused_level)

    def level(self, level: int, which: str = "A"):
        """(row_ptr, col_indices, values, nrows, ncols) of A_l, R_l or P_l."""
        w = {"A": 0, "R": 1, "P": 2}[which]
        n, m, z = C.c_int64(), C.c_int64(), C.c_int64()
        N.call("cfdla_amg_export_level", self.handle, level, w, C.byref(n), C.byref(m), C.byref(z), None, None,
               None)
        rp = np.empty(n.value + 1, np.int64)
        ci = np.empty(z.value, np.int64)
        v = np.empty(z.value)
        N.call("cfdla_amg_export_level", self.handle, level, w, C.byref(n), C.byref(m), C.byref(z), N.ptr(rp),
               N.ptr(ci), N.ptr(v))
        return rp, ci, v, n.value, m.value

    def level_map(self, level: int):
        n = self.statistics.level_sizes[level]
        out = np.empty(n, np.int64)
        N.call("cfdla_amg_export_map", self.handle, level, N.ptr(out))
        return out
