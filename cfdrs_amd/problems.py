"""Deterministic problem generators for the five workloads.

All random inputs come from SplitMix64 streams in libcfdla
(``cfdla_synth_uniform``), so the host, the GPU box and the tests see the
same matrices and right-hand sides.  Matrix assembly runs in the library:

* ``pressure_correction_2d``: cell-centred 5-point pressure-correction
  Laplacian, all cells fluid, east wall Dirichlet
  (crates/cfd-2d/src/pressure_velocity/correction.rs:273-343).
* ``varcoef``: cell-centred variable-coefficient Poisson with harmonic-mean
  face coefficients (2D or 3D).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _native as N

SEED_C1 = 1001
SEED_C2_K = 1002
SEED_C2_B = 10002
SEED_C5 = 1005


def uniform(seed: int, n: int, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    out = np.empty(n)
    N.call("cfdla_synth_uniform", C.c_uint64(seed), lo, hi, n, N.ptr(out))
    return out


def pressure_correction_2d(nxc: int, nyc: int, dx: float, dy: float, dirichlet=(0, 1, 0, 0)):
    """(row_ptr, col, values) of the pressure-correction matrix; without a
    Dirichlet side the first cell is pinned (nxc*nyc - 1 rows)."""
    d = np.asarray(dirichlet, np.int32)
    nnz = C.c_int64()
    N.call("cfdla_problem_pressure_correction", nxc, nyc, dx, dy, N.ptr(d), C.byref(nnz), None, None, None)
    rows = nxc * nyc - (0 if np.any(d) else 1)
    rp = np.empty(rows + 1, np.int64)
    ci = np.empty(nnz.value, np.int64)
    v = np.empty(nnz.value)
    N.call("cfdla_problem_pressure_correction", nxc, nyc, dx, dy, N.ptr(d), C.byref(nnz), N.ptr(rp), N.ptr(ci),
           N.ptr(v))
    return rp, ci, v


def simple_pressure(nx: int, ny: int, dx: float, dy: float, side, u, v, p, d_u, d_v, rho,
                    positive_diagonal: bool = False):
    """(row_ptr, col, values, rhs, max_residual) of the SIMPLE
    pressure-correction system on an nx x ny node grid
    (crates/cfd-2d/src/solvers/simple/pressure.rs:18-240); side = (west,
    east, south, north) of 0 plain / 1 pressure / 2 periodic.
    positive_diagonal negates the interior rows and their rhs (same
    solution; the form to precondition with AMG)."""
    sd = np.asarray(side, np.int32)
    f = [np.ascontiguousarray(a, np.float64) for a in (u, v, p, d_u, d_v, rho)]
    for a in f:
        if a.size != nx * ny:
            raise N.CfdlaError(N.ERR_DIMENSION, "fields must have nx * ny entries")
    nnz = C.c_int64()
    pos = 1 if positive_diagonal else 0
    N.call("cfdla_problem_simple_pressure", nx, ny, dx, dy, N.ptr(sd), pos, *(N.ptr(a) for a in f), C.byref(nnz),
           None, None, None, None, None)
    rp = np.empty(nx * ny + 1, np.int64)
    ci = np.empty(nnz.value, np.int64)
    val = np.empty(nnz.value)
    rhs = np.empty(nx * ny)
    mr = C.c_double()
    N.call("cfdla_problem_simple_pressure", nx, ny, dx, dy, N.ptr(sd), pos, *(N.ptr(a) for a in f), C.byref(nnz),
           N.ptr(rp), N.ptr(ci), N.ptr(val), N.ptr(rhs), C.byref(mr))
    return rp, ci, val, rhs, mr.value


def varcoef(nx: int, ny: int, nz: int, inv_h2: float, k, dir_mask: int = 0, dir_val=None):
    """(row_ptr, col, values, rhs_bc) of the variable-coefficient operator.
    dir_mask bits: 0 -x, 1 +x, 2 -y, 3 +y, 4 -z, 5 +z."""
    k = np.ascontiguousarray(k, np.float64)
    dv = np.zeros(6) if dir_val is None else np.ascontiguousarray(dir_val, np.float64)
    n = nx * ny * nz
    nnz = C.c_int64()
    N.call("cfdla_problem_varcoef", nx, ny, nz, inv_h2, N.ptr(k), dir_mask, N.ptr(dv), C.byref(nnz), None, None,
           None, None)
    rp = np.empty(n + 1, np.int64)
    ci = np.empty(nnz.value, np.int64)
    v = np.empty(nnz.value)
    bc = np.empty(n)
    N.call("cfdla_problem_varcoef", nx, ny, nz, inv_h2, N.ptr(k), dir_mask, N.ptr(dv), C.byref(nnz), N.ptr(rp),
           N.ptr(ci), N.ptr(v), N.ptr(bc))
    return rp, ci, v, bc


def config1(n: int = 128):
    """2D pressure-correction Poisson on n x n cells, h = 1/n, east Dirichlet; b ~ U[-1,1]."""
    h = 1.0 / n
    rp, ci, v = pressure_correction_2d(n, n, h, h, (0, 1, 0, 0))
    return rp, ci, v, uniform(SEED_C1, n * n)


def config2_k(n: int = 128):
    """k = exp(g), g ~ U[-ln 10, ln 10] per cell."""
    return np.exp(uniform(SEED_C2_K, n ** 3, -math.log(10.0), math.log(10.0)))


def grf(nx: int, ny: int, corr_len: float = 16.0, seed: int = SEED_C5) -> np.ndarray:
    """Gaussian random field, covariance exp(-r^2 / (2 corr_len^2)), sample mean 0, deviation 1."""
    out = np.empty(nx * ny)
    N.call("cfdla_synth_grf", nx, ny, corr_len, C.c_uint64(seed), N.ptr(out))
    return out


SEED_C3 = 1003


def cavity_velocity(n: int = 512, lid: float = 1.0, noise: float = 0.05, seed: int = SEED_C3):
    """Node velocities of the lid-driven cavity on n x n nodes (unit square):
    u = lid (y)^4, v = 0, plus noise * lid * U[-1, 1] on interior nodes; the
    walls carry no-slip values and the lid (top row) u = lid."""
    y = np.linspace(0.0, 1.0, n)
    u = np.tile((lid * y ** 4)[:, None], (1, n)).ravel()
    v = np.zeros(n * n)
    r = uniform(seed, 2 * n * n).reshape(2, n, n)
    interior = np.zeros((n, n), bool)
    interior[1:-1, 1:-1] = True
    u = u.reshape(n, n)
    v = v.reshape(n, n)
    u[interior] += noise * lid * r[0][interior]
    v[interior] += noise * lid * r[1][interior]
    u[0, :] = u[:, 0] = u[:, -1] = 0.0
    u[-1, :] = lid
    v[0, :] = v[-1, :] = v[:, 0] = v[:, -1] = 0.0
    return u.ravel(), v.ravel()


def momentum_u(n, dx, dy, dt, mu, rho, u, v):
    """(row_ptr, col, values) of the first-order-upwind x-momentum operator."""
    u = np.ascontiguousarray(u, np.float64)
    v = np.ascontiguousarray(v, np.float64)
    nnz = C.c_int64()
    N.call("cfdla_problem_momentum_u", n, n, dx, dy, dt, mu, rho, N.ptr(u), N.ptr(v), C.byref(nnz), None, None, None)
    rp = np.empty(n * n + 1, np.int64)
    ci = np.empty(nnz.value, np.int64)
    vals = np.empty(nnz.value)
    N.call("cfdla_problem_momentum_u", n, n, dx, dy, dt, mu, rho, N.ptr(u), N.ptr(v), C.byref(nnz), N.ptr(rp),
           N.ptr(ci), N.ptr(vals))
    return rp, ci, vals


def pressure_rhs(n, dx, dy, rho, dt, u, v, pin=True):
    out = np.empty((n - 2) * (n - 2) - (1 if pin else 0))
    u = np.ascontiguousarray(u, np.float64)  # referenced until the call returns
    v = np.ascontiguousarray(v, np.float64)
    N.call("cfdla_problem_pressure_rhs", n, n, dx, dy, rho, dt, N.ptr(u), N.ptr(v), 1 if pin else 0, N.ptr(out))
    return out


def pressure_rhs_device(nx, ny, dx, dy, rho, dt, u, v, rhs, pin=True) -> None:
    """pressure_rhs on CUDA float64 tensors (u, v of nx * ny nodes; rhs of
    the interior cells, one fewer when pinned), bit for bit the host's."""
    import torch

    cells = (nx - 2) * (ny - 2) - (1 if pin else 0)
    for t, m in ((u, nx * ny), (v, nx * ny), (rhs, cells)):
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
                and t.numel() == m):
            raise N.CfdlaError(N.ERR_INVALID_ARG, "pressure_rhs_device takes contiguous CUDA float64 tensors")
    from .sparse import bind_torch_stream

    bind_torch_stream()
    N.call("cfdla_pressure_rhs_device", nx, ny, dx, dy, rho, dt, u.data_ptr(), v.data_ptr(), 1 if pin else 0,
           rhs.data_ptr())


def config3(n: int = 512, re: float = 1000.0, dt: float = 0.05):
    """Lid-driven cavity at Re = 1000 on n x n nodes: the x-momentum system
    (right-hand side rho V / dt u_old) and the pinned pressure-correction
    system on the (n-2)^2 interior cells."""
    h = 1.0 / (n - 1)
    mu, rho = 1.0 / re, 1.0
    u, v = cavity_velocity(n)
    mrp, mci, mv = momentum_u(n, h, h, dt, mu, rho, u, v)
    b_m = rho * h * h / dt * u
    b_m[np.r_[0:n, (n - 1) * n:n * n]] = u[np.r_[0:n, (n - 1) * n:n * n]]
    for j in range(n):
        b_m[j * n] = u[j * n]
        b_m[j * n + n - 1] = u[j * n + n - 1]
    prp, pci, pv = pressure_correction_2d(n - 2, n - 2, h, h, (0, 0, 0, 0))
    b_p = pressure_rhs(n, h, h, rho, dt, u, v, pin=True)
    return (mrp, mci, mv, b_m), (prp, pci, pv, b_p)


def config2(n: int = 128):
    """3D variable-coefficient Poisson on n^3 cells, Dirichlet on z = 0 (value 0); b ~ U[-1,1]."""
    k = config2_k(n)
    rp, ci, v, _ = varcoef(n, n, n, float(n * n), k, 1 << 4, np.zeros(6))
    return rp, ci, v, uniform(SEED_C2_B, n ** 3)
