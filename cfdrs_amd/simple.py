"""SIMPLE pressure-correction step with a device-resident system.

Mirrors the pressure half of ``SimpleAlgorithm``
(crates/cfd-2d/src/solvers/simple/pressure.rs): ``assemble_pressure_correction``
(:18-450) builds the pattern once and afterwards only rewrites values and the
right-hand side, ``solve_pressure_correction`` (:452-485) solves with
BiCGSTAB, ``apply_corrections`` (:487-520) updates p, u and v.  Here the
refresh and the corrections run on the GPU (``cfdla_simple_pressure_refresh_device``,
``cfdla_simple_corrections_device``), so a SIMPLE iteration moves no matrix
over PCIe; an AMG preconditioner built on the system is refreshed on the
device too (``AlgebraicMultigrid.recompute_device``).

Boundary sides map to three kinds (pressure_velocity/boundary.rs:13-66):
pressure-anchor sides (Dirichlet, pressure inlet / outlet) give identity
rows, periodic sides couple to the partner node, every other condition
(walls, outflow, velocity inlets) averages the nodes across the side.
"""
from __future__ import annotations

import numpy as np

from . import _native as N
from .problems import simple_pressure
from .sparse import CsrMatrix, _is_torch, bind_torch_stream

SIDE_PLAIN, SIDE_PRESSURE, SIDE_PERIODIC = 0, 1, 2
_KINDS = {
    "wall": SIDE_PLAIN, "wall_no_slip": SIDE_PLAIN, "outflow": SIDE_PLAIN, "velocity_inlet": SIDE_PLAIN,
    "plain": SIDE_PLAIN, "dirichlet": SIDE_PRESSURE, "pressure_inlet": SIDE_PRESSURE,
    "pressure_outlet": SIDE_PRESSURE, "pressure": SIDE_PRESSURE, "periodic": SIDE_PERIODIC,
}


def sides(boundaries: dict) -> tuple:
    """(west, east, south, north) kinds from a {side: condition} map; every
    side must be present (momentum/boundary/consistency.rs:64-70)."""
    out = []
    for name in ("west", "east", "south", "north"):
        if name not in boundaries:
            raise N.CfdlaError(N.ERR_CONFIG, f"Missing required boundary: {name}")
        kind = boundaries[name]
        out.append(_KINDS[kind] if isinstance(kind, str) else int(kind))
    return tuple(out)


def assemble_pressure_correction(nx, ny, dx, dy, side, u, v, p, d_u, d_v, rho, positive_diagonal=False):
    """Host assembly: (CsrMatrix, rhs, max |mass imbalance|)."""
    rp, ci, val, rhs, mr = simple_pressure(nx, ny, dx, dy, side, u, v, p, d_u, d_v, rho, positive_diagonal)
    return CsrMatrix.from_parts(val, ci, rp, nx * ny, nx * ny), rhs, mr


def _dev(a, name):
    import torch

    if not (_is_torch(a) and a.is_cuda and a.dtype == torch.float64 and a.is_contiguous()):
        raise N.CfdlaError(N.ERR_INVALID_ARG, f"{name} must be a contiguous CUDA float64 tensor")
    return a.data_ptr()


class SimplePressureCorrection:
    """Pressure-correction system of one grid, assembled once on the host
    (the first-build branch) and refreshed on the device afterwards.

    ``positive_diagonal`` stores the interior rows negated (a_P = +sum
    a_nb, rhs negated; the solution is unchanged).  The reference's sign
    (a_P = -sum a_nb beside +1 boundary rows) suits its unpreconditioned
    BiCGSTAB; an AMG hierarchy needs rows of one sign, since its Galerkin
    product sums the rows of an aggregate."""

    def __init__(self, nx: int, ny: int, dx: float, dy: float, side, fields=None, positive_diagonal=False):
        self.nx, self.ny, self.dx, self.dy = nx, ny, dx, dy
        self.side = tuple(side)
        self.positive_diagonal = bool(positive_diagonal)
        n = nx * ny
        if fields is None:
            z = np.zeros(n)
            fields = (z, z, z, z, z, np.ones(n))
        self.matrix, self.rhs_host, self.max_residual = assemble_pressure_correction(
            nx, ny, dx, dy, self.side, *fields, positive_diagonal=self.positive_diagonal)

    def assemble_device(self, u, v, p, d_u, d_v, rho, rhs, want_max: bool = True):
        """Refresh the matrix values and write rhs (CUDA float64 tensors of
        nx*ny entries); returns max |mass imbalance| (synchronizes) or None."""
        n = self.nx * self.ny
        for a, name in ((u, "u"), (v, "v"), (p, "p"), (d_u, "d_u"), (d_v, "d_v"), (rho, "rho"), (rhs, "rhs")):
            if a.numel() != n:
                raise N.CfdlaError(N.ERR_DIMENSION, f"{name} must have nx * ny entries")
        bind_torch_stream()
        import ctypes as C

        mr = C.c_double()
        N.call("cfdla_simple_pressure_refresh_device", self.matrix.handle, self.nx, self.ny, self.dx, self.dy,
               1 if self.positive_diagonal else 0, _dev(u, "u"), _dev(v, "v"), _dev(p, "p"), _dev(d_u, "d_u"), _dev(d_v, "d_v"), _dev(rho, "rho"),
               _dev(rhs, "rhs"), C.byref(mr) if want_max else None)
        self.matrix._values = None
        return mr.value if want_max else None

    def apply_corrections_device(self, p_prime, d_u, d_v, p, u, v, relaxation: float) -> None:
        """p += relaxation p', u -= d_u dp'/dx dx, v -= d_v dp'/dy dy at
        interior nodes (pressure.rs:487-520)."""
        n = self.nx * self.ny
        for a, name in ((p_prime, "p_prime"), (d_u, "d_u"), (d_v, "d_v"), (p, "p"), (u, "u"), (v, "v")):
            if a.numel() != n:
                raise N.CfdlaError(N.ERR_DIMENSION, f"{name} must have nx * ny entries")
        bind_torch_stream()
        N.call("cfdla_simple_corrections_device", self.nx, self.ny, self.dx, self.dy, float(relaxation),
               _dev(p_prime, "p_prime"), _dev(d_u, "d_u"), _dev(d_v, "d_v"), _dev(p, "p"), _dev(u, "u"),
               _dev(v, "v"))
