"""Differentiable pressure solves: a torch.autograd.Function whose forward
pass is one AMG-PCG solve and whose backward pass is the adjoint AMG-PCG
solve plus the device coefficient-gradient kernel, all on the GPU.

Problem: the cell-centred variable-coefficient Poisson operator on nx*ny
cells (face coefficient 2 kP kN / (kP + kN) / h^2), k = exp(g), p = 1 on the
west wall and p = 0 on the east wall (Dirichlet faces add k / h^2 to the
diagonal, as the pressure-correction operator's Dirichlet side does,
crates/cfd-2d/src/pressure_velocity/correction.rs:300-331), no-flux north and
south walls, no source.  The objective is the outlet flux through the east
wall, J = sum over east cells of k p (the Dirichlet face flux k (p - 0) / h
times the face length h).

Device residency: the matrix keeps one sparsity pattern; each new g is
assembled on the GPU (cfdla_varcoef_refresh_device, bit for bit the host
generator's values) and the AMG hierarchy, built once from the first g, is
brought up to date by the device Galerkin recompute with fixed transfer
operators (the reference's AlgebraicMultigrid::recompute, amg.rs:179-214).
J stays a device tensor; the host only enqueues work and polls convergence.

No reference counterpart: the reference has no adjoint or autodiff code.
"""
from __future__ import annotations

import numpy as np

from . import _native as N
from . import problems as P
from .linear_solver import IterativeSolverConfig, Termination, cg_preconditioned
from .multigrid import AlgebraicMultigrid, AMGConfig
from .sparse import CsrMatrix, bind_torch_stream

DIR_MASK = 0b0011          # west and east faces
DIR_VAL = (1.0, 0.0, 0.0, 0.0, 0.0, 0.0)


class FluxProblem:
    """The outlet-flux problem on an nx x ny grid, with its device matrix,
    right-hand side and AMG hierarchy (created on the first solve)."""

    def __init__(self, nx: int, ny: int, tol: float = 1e-10, amg_config: AMGConfig | None = None):
        self.nx, self.ny = nx, ny
        self.n = nx * ny
        self.inv_h2 = float(nx * nx)
        self.tol = tol
        self.amg_config = amg_config or AMGConfig.smoothed_aggregation()
        self.east = np.arange(ny) * nx + (nx - 1)
        self._a = None
        self._amg = None
        self._rhs = None
        self._east_dev = None
        self._k = None  # the coefficient tensor the device matrix currently holds

    # ---- host helpers (the oracle-side checks use these) ------------------------
    def assemble(self, g: np.ndarray):
        """Host assembly: (k, CsrMatrix, rhs) for log-coefficients g."""
        k = np.exp(g)
        rp, ci, v, rhs = P.varcoef(self.nx, self.ny, 1, self.inv_h2, k, DIR_MASK, np.array(DIR_VAL))
        return k, CsrMatrix.from_parts(v, ci, rp, k.size, k.size), rhs

    def objective(self, k, p):
        return float(np.sum(k[self.east] * p[self.east]))

    def dJ_dp(self, k):
        d = np.zeros(k.size)
        d[self.east] = k[self.east]
        return d

    def dJ_dk_explicit(self, p):
        d = np.zeros(p.size)
        d[self.east] = p[self.east]
        return d

    # ---- device path ----------------------------------------------------------
    def set_coefficients(self, k) -> None:
        """Assemble A(k) and b(k) on the device from a float64 CUDA tensor k and
        bring the hierarchy up to date (built from this k on the first call)."""
        import torch

        if self._a is None:
            rp, ci, v, _ = P.varcoef(self.nx, self.ny, 1, self.inv_h2, np.ones(self.n), DIR_MASK, np.array(DIR_VAL))
            self._a = CsrMatrix.from_parts(v, ci, rp, self.n, self.n)
            self._rhs = torch.empty(self.n, dtype=torch.float64, device=k.device)
            self._east_dev = torch.from_numpy(self.east).to(k.device)
        self._a.varcoef_refresh_device(self.nx, self.ny, 1, self.inv_h2, k, DIR_MASK, DIR_VAL, self._rhs)
        if self._amg is None:
            self._amg = AlgebraicMultigrid(self._a, self.amg_config)
        else:
            self._amg.recompute_device(self._a)
        self._k = k

    def _solve(self, b, x) -> None:
        cfg = IterativeSolverConfig.new(self.tol).with_max_iterations(2000)
        rep = cg_preconditioned(self._a, b, self._amg, x, cfg)
        if rep.termination not in (Termination.Converged, Termination.InitialResidual):
            raise RuntimeError(f"pressure solve did not converge: {rep}")

    def forward(self, g):
        """(J, k, p) for a float64 CUDA tensor g: one device assembly and one
        AMG-PCG solve; J is a 0-dim device tensor."""
        import torch

        k = torch.exp(g.detach()).contiguous()
        self.set_coefficients(k)
        p = torch.zeros(self.n, dtype=torch.float64, device=g.device)
        self._solve(self._rhs, p)
        return (k[self._east_dev] * p[self._east_dev]).sum(), k, p

    def gradient(self, k, p):
        """(dJ/dg, lambda) from the adjoint solve A^T lambda = dJ/dp (A is
        symmetric: the forward hierarchy serves) and the coefficient-gradient
        kernel; re-assembles A(k) first if another k was set since."""
        import torch

        if self._k is None or self._k.data_ptr() != k.data_ptr():
            self.set_coefficients(k)
        rhs = torch.zeros_like(p)
        rhs[self._east_dev] = k[self._east_dev]
        lam = torch.zeros_like(p)
        self._solve(rhs, lam)
        dk = torch.zeros_like(p)
        dk[self._east_dev] = p[self._east_dev]
        out = torch.empty_like(p)
        dv = np.ascontiguousarray(DIR_VAL[:4], np.float64)
        bind_torch_stream()
        N.call("cfdla_coef_grad_poisson2d_device", self.nx, self.ny, self.inv_h2, DIR_MASK, N.ptr(dv),
               k.data_ptr(), p.data_ptr(), lam.data_ptr(), dk.data_ptr(), out.data_ptr())
        return out, lam


def flux_and_gradient(problem: FluxProblem, g, device: str = "cuda"):
    """J(g) (a float) and dJ/dg, p and lambda (device tensors) by one forward
    and one adjoint AMG-PCG solve on the GPU; g is a host array or a CUDA
    tensor."""
    import torch

    if not isinstance(g, torch.Tensor):
        g = torch.from_numpy(np.ascontiguousarray(g, np.float64)).to(device)
    J, k, p = problem.forward(g)
    grad, lam = problem.gradient(k, p)
    return float(J), grad, p, lam


def _autograd_function():
    import torch

    class OutletFlux(torch.autograd.Function):
        """J = outlet flux of the pressure field for log-coefficients g (a
        float64 CUDA tensor of nx*ny cells, x fastest).  Forward: one solve;
        backward: the adjoint solve and the gradient kernel."""

        @staticmethod
        def forward(ctx, g, problem: FluxProblem):
            if not g.is_cuda or g.dtype != torch.float64 or g.numel() != problem.n:
                raise N.CfdlaError(N.ERR_INVALID_ARG, "g must be a float64 CUDA tensor of nx*ny cells")
            J, k, p = problem.forward(g)
            ctx.problem = problem
            ctx.save_for_backward(k, p)
            return J

        @staticmethod
        def backward(ctx, grad_out):
            k, p = ctx.saved_tensors
            grad, _ = ctx.problem.gradient(k, p)
            return grad_out * grad, None

    return OutletFlux


_FN = None


def outlet_flux(g, problem: FluxProblem):
    """Differentiable J(g); the gradient comes from the adjoint solve."""
    global _FN
    if _FN is None:
        _FN = _autograd_function()
    return _FN.apply(g, problem)
