"""Differentiable pressure solves: a torch.autograd.Function whose backward is
an adjoint AMG-PCG solve plus the device coefficient-gradient kernel.

Problem: the cell-centred variable-coefficient Poisson operator on nx*ny
cells (face coefficient 2 kP kN / (kP + kN) / h^2), k = exp(g), p = 1 on the
west wall and p = 0 on the east wall (Dirichlet faces add k / h^2 to the
diagonal, as the pressure-correction operator's Dirichlet side does,
crates/cfd-2d/src/pressure_velocity/correction.rs:300-331), no-flux north and
south walls, no source.  The objective is the outlet flux through the east
wall, J = sum over east cells of k p (the Dirichlet face flux k (p - 0) / h
times the face length h).

No reference counterpart: the reference has no adjoint or autodiff code.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from . import problems as P
from .linear_solver import IterativeSolverConfig, Termination, cg_preconditioned
from .multigrid import AlgebraicMultigrid, AMGConfig
from .sparse import CsrMatrix, bind_torch_stream

DIR_MASK = 0b0011          # west and east faces
DIR_VAL = (1.0, 0.0, 0.0, 0.0, 0.0, 0.0)


class FluxProblem:
    """Assembly and solves for one log-coefficient field g (host arrays)."""

    def __init__(self, nx: int, ny: int, tol: float = 1e-10, amg_config: AMGConfig | None = None):
        self.nx, self.ny = nx, ny
        self.inv_h2 = float(nx * nx)
        self.tol = tol
        self.amg_config = amg_config or AMGConfig.smoothed_aggregation()

    def assemble(self, g: np.ndarray):
        k = np.exp(g)
        rp, ci, v, rhs = P.varcoef(self.nx, self.ny, 1, self.inv_h2, k, DIR_MASK, np.array(DIR_VAL))
        return k, CsrMatrix.from_parts(v, ci, rp, k.size, k.size), rhs

    def objective(self, k, p):
        east = np.arange(self.ny) * self.nx + (self.nx - 1)
        return float(np.sum(k[east] * p[east]))

    def dJ_dp(self, k):
        d = np.zeros(k.size)
        east = np.arange(self.ny) * self.nx + (self.nx - 1)
        d[east] = k[east]
        return d

    def dJ_dk_explicit(self, p):
        d = np.zeros(p.size)
        east = np.arange(self.ny) * self.nx + (self.nx - 1)
        d[east] = p[east]
        return d


def _solve(a: CsrMatrix, amg, b, x, tol):
    cfg = IterativeSolverConfig.new(tol).with_max_iterations(2000)
    rep = cg_preconditioned(a, b, amg, x, cfg)
    if rep.termination not in (Termination.Converged, Termination.InitialResidual):
        raise RuntimeError(f"pressure solve did not converge: {rep}")
    return rep


def flux_and_gradient(problem: FluxProblem, g: np.ndarray, device: str = "cuda"):
    """J(g) and dJ/dg by one forward and one adjoint AMG-PCG solve on the GPU."""
    import torch

    k, a, rhs = problem.assemble(np.ascontiguousarray(g, np.float64))
    amg = AlgebraicMultigrid(a, problem.amg_config)
    dev = torch.device(device)
    b = torch.from_numpy(rhs).to(dev)
    p = torch.zeros(k.size, dtype=torch.float64, device=dev)
    _solve(a, amg, b, p, problem.tol)
    p_host = p.cpu().numpy()
    J = problem.objective(k, p_host)
    # adjoint: A^T lambda = dJ/dp (A symmetric: the forward hierarchy is reused)
    lam = torch.zeros_like(p)
    gl = torch.from_numpy(problem.dJ_dp(k)).to(dev)
    _solve(a, amg, gl, lam, problem.tol)
    kd = torch.from_numpy(k).to(dev)
    dk = torch.from_numpy(problem.dJ_dk_explicit(p_host)).to(dev)
    out = torch.empty_like(p)
    dv = np.ascontiguousarray(DIR_VAL[:4], np.float64)
    bind_torch_stream()
    N.call("cfdla_coef_grad_poisson2d_device", problem.nx, problem.ny, problem.inv_h2, DIR_MASK, N.ptr(dv),
           kd.data_ptr(), p.data_ptr(), lam.data_ptr(), dk.data_ptr(), out.data_ptr())
    return J, out, p, lam


def _autograd_function():
    import torch

    class OutletFlux(torch.autograd.Function):
        """J = outlet flux of the pressure field for log-coefficients g
        (a float64 tensor of nx*ny cells, x fastest)."""

        @staticmethod
        def forward(ctx, g, problem: FluxProblem):
            J, grad, _, _ = flux_and_gradient(problem, g.detach().cpu().numpy(), device=str(g.device)
                                              if g.is_cuda else "cuda")
            ctx.save_for_backward(grad.to(g.device))
            return torch.tensor(J, dtype=g.dtype, device=g.device)

        @staticmethod
        def backward(ctx, grad_out):
            (grad,) = ctx.saved_tensors
            return grad_out * grad, None

    return OutletFlux


_FN = None


def outlet_flux(g, problem: FluxProblem):
    """Differentiable J(g); the gradient comes from the adjoint solve."""
    global _FN
    if _FN is None:
        _FN = _autograd_function()
    return _FN.apply(g, problem)
