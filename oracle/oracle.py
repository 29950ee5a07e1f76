"""ctypes wrapper of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: this module is the checker.  It may be imported by
tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
and by nothing else.  The product package (cfdrs_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

I64P = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
U8P = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")

# termination codes (mirror cfdla.h CFDLA_TERM_*)
T_CONVERGED, T_INITIAL, T_MAXIT, T_BREAKDOWN, T_NONPOS = 0, 1, 2, 3, 4
# smoother kinds
SM_GS, SM_SGS, SM_JACOBI, SM_SOR, SM_CHEBYSHEV = 0, 1, 2, 3, 4


class Report(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("initial_residual_norm", C.c_double),
                ("final_residual_norm", C.c_double), ("threshold", C.c_double),
                ("termination", C.c_int64)]


class AmgConfig(C.Structure):
    _fields_ = [("max_levels", C.c_int64), ("min_coarse", C.c_int64),
                ("coarsening", C.c_int32), ("interp", C.c_int32),
                ("smoother", C.c_int32), ("coarse_solve", C.c_int32),
                ("pre", C.c_int64), ("post", C.c_int64),
                ("omega", C.c_double), ("theta", C.c_double), ("sa_omega", C.c_double),
                ("max_agg", C.c_int64)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)


def set_threads(n: int) -> None:
    """OpenMP threads for the solve-phase loops (1 = serial, index-order sums)."""
    lib().orc_set_threads(int(n))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        sig = {
            "orc_csr_new": (vp, [C.c_int64, C.c_int64, I64P, I64P, F64P]),
            "orc_csr_free": (None, [vp]),
            "orc_csr_info": (None, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
            "orc_csr_export": (None, [vp, I64P, I64P, F64P]),
            "orc_spmv": (None, [vp, F64P, F64P]),
            "orc_transpose": (vp, [vp]),
            "orc_spgemm": (vp, [vp, vp]),
            "orc_strength": (vp, [vp, C.c_double]),
            "orc_coarsen": (C.c_int64, [vp, C.c_int, C.c_double, C.c_int64, I64P, I64P]),
            "orc_direct_interp": (vp, [I64P, C.c_int64, C.c_int64]),
            "orc_classical_interp": (vp, [vp, I64P, C.c_int64, vp]),
            "orc_smooth_interp": (vp, [vp, vp, C.c_double]),
            "orc_smoother": (None, [vp, C.c_int, C.c_double, F64P, F64P, C.c_int64]),
            "orc_chebyshev_bounds": (None, [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "orc_amg_new": (vp, [vp, C.POINTER(AmgConfig)]),
            "orc_amg_free": (None, [vp]),
            "orc_amg_num_levels": (C.c_int64, [vp]),
            "orc_amg_complexity": (None, [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "orc_amg_level": (vp, [vp, C.c_int64, C.c_int]),
            "orc_amg_level_map": (None, [vp, C.c_int64, I64P]),
            "orc_amg_apply": (None, [vp, F64P, F64P]),
            "orc_solve": (None, [vp, C.c_int, C.c_int, vp, F64P, F64P, C.c_double, C.c_double,
                                 C.c_int64, C.c_int64, C.POINTER(Report)]),
            "orc_restart_ladder": (C.c_int64, [C.c_int64]),
            "orc_set_threads": (None, [C.c_int32]),
            "orc_jacobi_inv_diag": (None, [vp, F64P]),
            "orc_dense_solve": (C.c_int, [vp, F64P, F64P]),
            "orc_assemble_pressure_correction": (vp, [C.c_int64, C.c_int64, C.c_double, C.c_double, I32P]),
            "orc_assemble_varcoef": (vp, [C.c_int64, C.c_int64, C.c_int64, C.c_double, F64P, C.c_int32,
                                          F64P, F64P]),
            "orc_assemble_momentum_u": (vp, [C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double,
                                             C.c_double, C.c_double, F64P, F64P]),
            "orc_assemble_simple_pressure": (vp, [C.c_int64, C.c_int64, C.c_double, C.c_double, I32P, F64P, F64P,
                                                  F64P, F64P, F64P, F64P, F64P, F64P]),
            "orc_simple_corrections": (None, [C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double, F64P, F64P,
                                              F64P, F64P, F64P, F64P]),
            "orc_assemble_network": (vp, [C.c_int64, C.c_int64, I64P, I64P, F64P, U8P, F64P, F64P, F64P]),
            "orc_detect_spd": (C.c_int, [vp]),
            "orc_poiseuille_number": (C.c_double, [C.c_double]),
            "orc_rect_resistance": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_double]),
            "orc_network_cascade": (C.c_int, [vp, F64P, F64P, C.c_double, C.c_int64, C.c_int,
                                              C.POINTER(C.c_int64)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class Csr:
    """Owning handle of an oracle CSR matrix (int64 indices, f64 values)."""

    def __init__(self, handle):
        if not handle:
            raise ValueError("oracle returned no matrix")
        self.h = handle

    @classmethod
    def from_arrays(cls, rp, ci, v, n, m=None):
        m = n if m is None else m
        return cls(lib().orc_csr_new(n, m, np.ascontiguousarray(rp, np.int64),
                                     np.ascontiguousarray(ci, np.int64),
                                     np.ascontiguousarray(v, np.float64)))

    @classmethod
    def from_dense(cls, d):
        d = np.asarray(d, np.float64)
        rp, ci, v = [0], [], []
        for row in d:
            nz = np.nonzero(row)[0]
            ci.extend(nz.tolist())
            v.extend(row[nz].tolist())
            rp.append(len(ci))
        return cls.from_arrays(np.array(rp), np.array(ci, np.int64), np.array(v), d.shape[0], d.shape[1])

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_csr_free(self.h)
            self.h = None

    @property
    def shape(self):
        n, m, z = C.c_int64(), C.c_int64(), C.c_int64()
        lib().orc_csr_info(self.h, C.byref(n), C.byref(m), C.byref(z))
        return n.value, m.value, z.value

    def arrays(self):
        n, m, nnz = self.shape
        rp = np.empty(n + 1, np.int64)
        ci = np.empty(nnz, np.int64)
        v = np.empty(nnz, np.float64)
        lib().orc_csr_export(self.h, rp, ci, v)
        return rp, ci, v

    def dense(self):
        n, m, _ = self.shape
        rp, ci, v = self.arrays()
        d = np.zeros((n, m))
        for i in range(n):
            for k in range(rp[i], rp[i + 1]):
                d[i, ci[k]] += v[k]
        return d

    def spmv(self, x):
        n, m, _ = self.shape
        y = np.empty(n)
        lib().orc_spmv(self.h, np.ascontiguousarray(x, np.float64), y)
        return y

    def transpose(self):
        return Csr(lib().orc_transpose(self.h))

    def __matmul__(self, other):
        return Csr(lib().orc_spgemm(self.h, other.h))


def strength(a: Csr, theta: float) -> Csr:
    return Csr(lib().orc_strength(a.h, theta))


def coarsen(a: Csr, kind: str, theta: float = 0.25, max_size: int = 8):
    n = a.shape[0]
    f2c = np.empty(n, np.int64)
    cp = np.empty(n, np.int64)
    code = {"rs": 0, "aggregation": 1, "hybrid": 2, "three_phase": 3}[kind]
    nc = lib().orc_coarsen(a.h, code, theta, max_size, f2c, cp)
    return cp[:nc].copy(), f2c


def direct_interp(f2c, n, nc) -> Csr:
    return Csr(lib().orc_direct_interp(np.ascontiguousarray(f2c, np.int64), n, nc))


def classical_interp(a: Csr, cpts, s: Csr) -> Csr:
    cp = np.ascontiguousarray(cpts, np.int64)
    return Csr(lib().orc_classical_interp(a.h, cp, len(cp), s.h))


def smooth_interp(a: Csr, pt: Csr, omega: float) -> Csr:
    return Csr(lib().orc_smooth_interp(a.h, pt.h, omega))


def smoother(a: Csr, kind: int, omega: float, x, b, iters: int):
    x = np.array(x, np.float64)
    lib().orc_smoother(a.h, kind, omega, x, np.ascontiguousarray(b, np.float64), iters)
    return x


def chebyshev_bounds(a: Csr):
    lo, hi = C.c_double(), C.c_double()
    lib().orc_chebyshev_bounds(a.h, C.byref(lo), C.byref(hi))
    return lo.value, hi.value


COARSENING = {"rs": 0, "aggregation": 1, "hybrid": 2, "three_phase": 3}
INTERP = {"classical": 0, "direct": 1, "smoothed": 2}
SMOOTHER = {"gauss_seidel": 0, "sgs": 1, "jacobi": 2, "sor": 3, "chebyshev": 4}


class Amg:
    """Oracle AMG hierarchy (multigrid/amg.rs restated)."""

    def __init__(self, a: Csr, *, max_levels=10, min_coarse=50, coarsening="rs", interp="classical",
                 smoother="gauss_seidel", pre=2, post=2, omega=1.0, theta=0.25, coarse_solve="sweeps",
                 sa_omega=2.0 / 3.0, max_agg=8):
        cfg = AmgConfig(max_levels, min_coarse, COARSENING[coarsening], INTERP[interp], SMOOTHER[smoother],
                        {"sweeps": 0, "direct": 1}[coarse_solve], pre, post, omega, theta, sa_omega, max_agg)
        self.h = lib().orc_amg_new(a.h, C.byref(cfg))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_amg_free(self.h)
            self.h = None

    @property
    def num_levels(self):
        return lib().orc_amg_num_levels(self.h)

    def complexity(self):
        op, gr = C.c_double(), C.c_double()
        lib().orc_amg_complexity(self.h, C.byref(op), C.byref(gr))
        return op.value, gr.value

    def level(self, l, which="A") -> Csr:
        return Csr(lib().orc_amg_level(self.h, l, {"A": 0, "R": 1, "P": 2}[which]))

    def level_map(self, l):
        n = self.level(l).shape[0]
        out = np.empty(n, np.int64)
        lib().orc_amg_level_map(self.h, l, out)
        return out

    def apply(self, r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.empty_like(r)
        lib().orc_amg_apply(self.h, r, z)
        return z


def solve(a: Csr, b, x0=None, *, kind="cg", prec="identity", amg: Amg | None = None, tol_abs=0.0,
          tol_rel=1e-10, max_it=1000, restart=30):
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros_like(b) if x0 is None else np.array(x0, np.float64)
    rep = Report()
    k = {"cg": 0, "bicgstab": 1, "gmres": 2}[kind]
    p = {"identity": 0, "jacobi": 1, "amg": 2}[prec]
    lib().orc_solve(a.h, k, p, amg.h if amg is not None else None, b, x, tol_abs, tol_rel, max_it, restart,
                    C.byref(rep))
    return x, rep


def restart_ladder(m: int) -> int:
    return lib().orc_restart_ladder(m)


def jacobi_inv_diag(a: Csr):
    out = np.empty(a.shape[0])
    lib().orc_jacobi_inv_diag(a.h, out)
    return out


def assemble_pressure_correction(nxc, nyc, dx, dy, dirichlet=(0, 1, 0, 0)) -> Csr:
    return Csr(lib().orc_assemble_pressure_correction(nxc, nyc, dx, dy, np.array(dirichlet, np.int32)))


def assemble_varcoef(nx, ny, nz, inv_h2, k, dir_mask, dir_val):
    rhs = np.empty(nx * ny * nz)
    a = Csr(lib().orc_assemble_varcoef(nx, ny, nz, inv_h2, np.ascontiguousarray(k, np.float64), dir_mask,
                                       np.ascontiguousarray(dir_val, np.float64), rhs))
    return a, rhs


def assemble_momentum_u(nx, ny, dx, dy, dt, mu, rho, u, v) -> Csr:
    return Csr(lib().orc_assemble_momentum_u(nx, ny, dx, dy, dt, mu, rho, np.ascontiguousarray(u, np.float64),
                                             np.ascontiguousarray(v, np.float64)))


def assemble_simple_pressure(nx, ny, dx, dy, side, u, v, p, du, dv, rho):
    """SIMPLE pressure-correction system: (matrix, rhs, max |mass imbalance|)."""
    f = [np.ascontiguousarray(a, np.float64) for a in (u, v, p, du, dv, rho)]
    for a in f:
        assert a.size == nx * ny
    rhs = np.empty(nx * ny)
    mr = np.zeros(1)
    a = Csr(lib().orc_assemble_simple_pressure(nx, ny, dx, dy, np.array(side, np.int32), *f, rhs, mr))
    return a, rhs, float(mr[0])


def simple_corrections(nx, ny, dx, dy, relax, pp, du, dv, p, u, v):
    """In-place p, u, v corrections after the SIMPLE pressure solve."""
    for a in (p, u, v):
        assert a.dtype == np.float64 and a.flags.c_contiguous and a.size == nx * ny
    lib().orc_simple_corrections(nx, ny, dx, dy, relax, np.ascontiguousarray(pp, np.float64),
                                 np.ascontiguousarray(du, np.float64), np.ascontiguousarray(dv, np.float64), p, u, v)


def assemble_network(n, ei, ej, g, is_dir, dir_val, neumann=None):
    rhs = np.empty(n)
    neu = np.zeros(n) if neumann is None else np.ascontiguousarray(neumann, np.float64)
    h = lib().orc_assemble_network(n, len(ei), np.ascontiguousarray(ei, np.int64), np.ascontiguousarray(ej, np.int64),
                                   np.ascontiguousarray(g, np.float64), np.ascontiguousarray(is_dir, np.uint8),
                                   np.ascontiguousarray(dir_val, np.float64), neu, rhs)
    if not h:
        raise ValueError("invalid network (self-loop or non-positive conductance)")
    return Csr(h), rhs


def detect_spd(a: Csr) -> bool:
    return bool(lib().orc_detect_spd(a.h))


def poiseuille_number(eps: float) -> float:
    return lib().orc_poiseuille_number(eps)


def rect_resistance(w, h, L, mu) -> float:
    return lib().orc_rect_resistance(w, h, L, mu)


def network_cascade(a: Csr, b, tol=1e-6, max_it=1000, use_cg=True):
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros_like(b)
    it = C.c_int64()
    path = lib().orc_network_cascade(a.h, b, x, tol, max_it, int(use_cg), C.byref(it))
    return x, path, it.value


def coef_grad_2d(nx, ny, inv_h2, dir_mask, dir_val, k, p, lam, dJdk=None):
    """Adjoint dJ/dg (k = exp g) of the 2D variable-coefficient operator."""
    L = lib()
    fn = L.orc_coef_grad_2d
    fn.restype = None
    fn.argtypes = [C.c_int64, C.c_int64, C.c_double, C.c_int32, F64P, F64P, F64P, F64P, C.c_void_p, F64P]
    out = np.empty(nx * ny)
    dj = None if dJdk is None else np.ascontiguousarray(dJdk, np.float64)
    fn(nx, ny, inv_h2, dir_mask, np.ascontiguousarray(dir_val, np.float64), np.ascontiguousarray(k, np.float64),
       np.ascontiguousarray(p, np.float64), np.ascontiguousarray(lam, np.float64),
       None if dj is None else dj.ctypes.data, out)
    return out


def rect_conductance_complex_step(w, h, L, mu, step: float = 1e-30):
    """G = 1/R for rectangular channels (crates/cfd-1d/src/physics/resistance/
    models/rectangular.rs:95-157,209-245: R = Po(eps) mu L / (2 A Dh^2),
    Dh = 4 A / (2 (w + h)), Po = 96 / ((1 + eps)^2 (1 - 192 eps / pi^5
    tanh(pi / (2 eps)))), eps = min(w, h) / max(w, h)) and dG/dw by the
    complex-step derivative (Squire & Trapp, SIAM Rev. 40 (1998) 110):
    dG/dw = Im G(w + i s) / s, exact to rounding for this analytic formula
    and independent of any hand-derived chain rule.  Arrays of equal length."""
    w = np.asarray(w, np.float64)
    h = np.broadcast_to(np.asarray(h, np.float64), w.shape)
    L = np.broadcast_to(np.asarray(L, np.float64), w.shape)

    def resistance(wc):
        area = wc * h
        dh = 4.0 * area / (2.0 * (wc + h))
        narrow = wc.real <= h
        eps = np.where(narrow, wc / h, h / wc)
        po = 96.0 / ((1.0 + eps) ** 2 * (1.0 - (192.0 * eps / np.pi ** 5) * np.tanh(np.pi / (2.0 * eps))))
        return po * mu * L / (2.0 * area * dh * dh)

    g = 1.0 / resistance(w.astype(np.complex128))
    s = step * w
    dg = (1.0 / resistance(w + 1j * s)).imag / s
    return g.real.copy(), dg


def network_outlet_gradient(n, ei, ej, width, height, length, mu, is_dir, dir_val, outlet):
    """Outlet flux of one network and its gradient in the channel widths.

    J = sum over edges e = (i, j) of G_e [m_j (p_i - p_j) + m_i (p_j - p_i)]
    (the flow into the nodes with outlet mask m = 1, with flows G_e (p_i - p_j)
    as the reference's update_from_solution forms them for linear channels,
    crates/cfd-1d/src/domain/network/wrapper.rs:349-428), p from the oracle's
    assembly (matrix_assembly.rs:163-247) and a dense FP64 solve.  Adjoint:
    A^T lambda = dJ/dp (zero on Dirichlet rows), then
    dJ/dG_e = dJ/dG_e|explicit - (lambda_i - lambda_j)(p_i - p_j) and
    dJ/dw_e = dJ/dG_e dG_e/dw_e (complex step).  Returns (J, dJ/dw, p)."""
    ei = np.asarray(ei, np.int64)
    ej = np.asarray(ej, np.int64)
    g, dg = rect_conductance_complex_step(width, height, length, mu)
    a, rhs = assemble_network(n, ei, ej, g, is_dir, dir_val)
    A = a.dense()
    p = np.linalg.solve(A, rhs)
    m = np.asarray(outlet, np.float64)
    explicit = m[ej] * (p[ei] - p[ej]) + m[ei] * (p[ej] - p[ei])
    J = float(np.sum(g * explicit))
    dJdp = np.zeros(n)
    np.add.at(dJdp, ei, g * (m[ej] - m[ei]))
    np.add.at(dJdp, ej, g * (m[ei] - m[ej]))
    dJdp[np.asarray(is_dir, bool)] = 0.0
    lam = np.linalg.solve(A.T, dJdp)
    dJdG = explicit - (lam[ei] - lam[ej]) * (p[ei] - p[ej])
    return J, dJdG * dg, p
