"""Batched 1D Kirchhoff networks on the GPU.

Each network is assembled as ``MatrixAssembler::assemble_into``
(crates/cfd-1d/src/solver/core/matrix_assembly.rs:163-247) and solved with
Jacobi-PCG to the row-equilibrated residual target of
``LinearSystemSolver`` (crates/cfd-1d/src/solver/core/linear_system.rs:114-246).
Thousands of networks go to one persistent launch; the row-bin schedule
(which networks a single wavefront solves, how many resident blocks per
compute unit) is chosen by ``NetworkBatch.autotune`` and cached as JSON.

Rectangular-duct conductances follow
crates/cfd-1d/src/physics/resistance/models/rectangular.rs:95-157,209-245.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .linear_solver import IterativeSolverConfig, Precision, SolveReport, Termination
from .sparse import _is_torch, bind_torch_stream

WATER_VISCOSITY = 1.002e-3  # Pa s
CHANNEL_HEIGHT = 50e-6      # m
INLET_PRESSURE = 1000.0     # Pa


@dataclass
class SynthNetworks:
    node_offset: np.ndarray
    edge_offset: np.ndarray
    edge_i: np.ndarray
    edge_j: np.ndarray
    width: np.ndarray
    length: np.ndarray

    @property
    def count(self) -> int:
        return self.node_offset.size - 1

    def sizes(self) -> np.ndarray:
        return np.diff(self.node_offset)


def synth_networks(count: int = 4096, seed_base: int = 4000, nmin: int = 50, nmax: int = 2000) -> SynthNetworks:
    """Planar channel networks: random spanning tree of a grid plus extra grid
    edges (degrees 2-4); graph g is seeded with seed_base + g."""
    no = np.empty(count + 1, np.int64)
    eo = np.empty(count + 1, np.int64)
    N.call("cfdla_synth_networks", count, C.c_uint64(seed_base), nmin, nmax, N.ptr(no), N.ptr(eo), None, None, None,
           None)
    E = int(eo[-1])
    ei, ej = np.empty(E, np.int64), np.empty(E, np.int64)
    w, L = np.empty(E), np.empty(E)
    N.call("cfdla_synth_networks", count, C.c_uint64(seed_base), nmin, nmax, N.ptr(no), N.ptr(eo), N.ptr(ei),
           N.ptr(ej), N.ptr(w), N.ptr(L))
    return SynthNetworks(no, eo, ei, ej, w, L)


def rect_conductance(width, height, length, viscosity=WATER_VISCOSITY):
    """(G, dG/dw) of rectangular ducts, G = 1 / R."""
    w = np.ascontiguousarray(width, np.float64)
    h = np.ascontiguousarray(np.broadcast_to(height, w.shape), np.float64)
    L = np.ascontiguousarray(np.broadcast_to(length, w.shape), np.float64)
    g, dg = np.empty_like(w), np.empty_like(w)
    N.call("cfdla_rect_conductance", w.size, N.ptr(w), N.ptr(h), N.ptr(L), viscosity, N.ptr(g), N.ptr(dg))
    return g, dg


def rect_conductance_device(width, height, length, viscosity=WATER_VISCOSITY):
    """(G, dG/dw) of rectangular ducts on float64 CUDA tensors (same shape)."""
    import torch

    w = width.contiguous()
    h = torch.broadcast_to(torch.as_tensor(height, dtype=torch.float64, device=w.device), w.shape).contiguous()
    L = torch.broadcast_to(torch.as_tensor(length, dtype=torch.float64, device=w.device), w.shape).contiguous()
    if w.dtype != torch.float64 or not w.is_cuda:
        raise N.CfdlaError(N.ERR_INVALID_ARG, "width must be a float64 CUDA tensor")
    g, dg = torch.empty_like(w), torch.empty_like(w)
    bind_torch_stream()
    N.call("cfdla_rect_conductance_device", w.numel(), w.data_ptr(), h.data_ptr(), L.data_ptr(), viscosity,
           g.data_ptr(), dg.data_ptr())
    return g, dg


def _require_cuda_f64(t, n: int, name: str) -> None:
    import torch

    if not (_is_torch(t) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and t.numel() == n):
        raise N.CfdlaError(N.ERR_INVALID_ARG, f"{name} must be a contiguous float64 CUDA tensor of {n}")


def _report(r: N.Report) -> SolveReport:
    return SolveReport(r.iterations, r.initial_residual_norm, r.final_residual_norm, r.threshold,
                       Termination(r.termination), r.outer_iterations, Precision(r.precision))


class NetworkBatch:
    """Independent networks packed into one block-diagonal system."""

    def __init__(self, node_offset, edge_offset, edge_i, edge_j, conductance, is_dirichlet, dirichlet_value,
                 neumann=None):
        self.node_offset = np.ascontiguousarray(node_offset, np.int64)
        self.edge_offset = np.ascontiguousarray(edge_offset, np.int64)
        self.edge_i = np.ascontiguousarray(edge_i, np.int64)
        self.edge_j = np.ascontiguousarray(edge_j, np.int64)
        self._g_host = np.ascontiguousarray(conductance, np.float64)
        self._g_dev = None  # device tensor of the latest device update, copied to the host on demand
        self.is_dirichlet = np.ascontiguousarray(is_dirichlet, np.uint8)
        self.dirichlet_value = np.ascontiguousarray(dirichlet_value, np.float64)
        self.neumann = None if neumann is None else np.ascontiguousarray(neumann, np.float64)
        h = C.c_void_p()
        N.call("cfdla_batch_create", self.node_offset.size - 1, N.ptr(self.node_offset), N.ptr(self.edge_offset),
               N.ptr(self.edge_i), N.ptr(self.edge_j), N.ptr(self._g_host), N.ptr(self.is_dirichlet),
               N.ptr(self.dirichlet_value), N.ptr(self.neumann), C.byref(h))
        self._h = h
        ns, nr, nz = C.c_int64(), C.c_int64(), C.c_int64()
        N.call("cfdla_batch_info", h, C.byref(ns), C.byref(nr), C.byref(nz))
        self.n_sys, self.n_rows, self.nnz = ns.value, nr.value, nz.value

    @property
    def handle(self):
        if self._h is None:
            raise ValueError("batch has been closed")
        return self._h

    @property
    def conductance(self) -> np.ndarray:
        """Edge conductances of the assembled systems (host copy)."""
        if self._g_dev is not None:
            self._g_host = self._g_dev.detach().cpu().numpy().copy()
            self._g_dev = None
        return self._g_host

    def update_conductance(self, conductance) -> None:
        g = np.ascontiguousarray(conductance, np.float64)
        N.call("cfdla_batch_update_conductance", self.handle, N.ptr(g))
        self._g_host, self._g_dev = g, None

    def update_conductance_device(self, conductance) -> None:
        """Re-assemble every network on the GPU from a float64 CUDA tensor of
        edge conductances (bit for bit the host assembly); no host copy."""
        import torch

        if not (_is_torch(conductance) and conductance.is_cuda and conductance.dtype == torch.float64
                and conductance.is_contiguous() and conductance.numel() == self.edge_i.size):
            raise N.CfdlaError(N.ERR_INVALID_ARG, "conductance must be a contiguous float64 CUDA tensor of n_edges")
        bind_torch_stream()
        N.call("cfdla_batch_update_conductance_device", self.handle, conductance.data_ptr())
        self._g_dev = conductance

    def export(self):
        """(row_ptr, col_indices, values, rhs) of the assembled batch."""
        rp = np.empty(self.n_rows + 1, np.int64)
        ci = np.empty(self.nnz, np.int64)
        v = np.empty(self.nnz)
        b = np.empty(self.n_rows)
        N.call("cfdla_batch_export", self.handle, N.ptr(rp), N.ptr(ci), N.ptr(v), N.ptr(b))
        return rp, ci, v, b

    @property
    def schedule(self) -> dict:
        s = N.BatchSchedule()
        N.call("cfdla_batch_get_schedule", self.handle, C.byref(s))
        return {"wave_max_rows": s.wave_max_rows, "blocks_per_cu": s.blocks_per_cu}

    @schedule.setter
    def schedule(self, d: dict) -> None:
        s = N.BatchSchedule(int(d["wave_max_rows"]), int(d["blocks_per_cu"]))
        N.call("cfdla_batch_set_schedule", self.handle, C.byref(s))

    def solve(self, x=None, config: IterativeSolverConfig | None = None, rhs=None):
        """Solve all networks.  x: numpy (host) or torch CUDA float64 packed
        pressures (initial guess in, solution out).  Returns (x, reports)."""
        config = config or IterativeSolverConfig.new(1e-10).with_max_iterations(20000)
        cfg = config.native(N.KIND_CG)
        reps = (N.Report * max(1, self.n_sys))()
        if x is not None and _is_torch(x):
            import torch

            if x.dtype != torch.float64 or x.numel() != self.n_rows or not x.is_contiguous():
                raise N.CfdlaError(N.ERR_DIMENSION, "x must be a contiguous float64 tensor of n_rows")
            bind_torch_stream()
            N.call("cfdla_batch_solve_device", self.handle, C.byref(cfg), None if rhs is None else rhs.data_ptr(),
                   x.data_ptr(), reps)
        else:
            if rhs is not None:
                raise N.CfdlaError(N.ERR_INVALID_ARG, "a right-hand side override takes device tensors")
            x = np.zeros(self.n_rows) if x is None else x
            if x.dtype != np.float64 or x.shape != (self.n_rows,) or not x.flags.c_contiguous:
                raise N.CfdlaError(N.ERR_DIMENSION, "x must be a contiguous float64 array of n_rows")
            N.call("cfdla_batch_solve", self.handle, C.byref(cfg), N.ptr(x), reps)
        return x, [_report(reps[k]) for k in range(self.n_sys)]

    def edge_gradient(self, p, lam, explicit_dg=None, dg_dparam=None):
        """Device tensors: dJ/dparam per edge from pressures p and adjoint lam."""
        import torch

        out = torch.empty(self.edge_i.size, dtype=torch.float64, device=p.device)
        bind_torch_stream()
        N.call("cfdla_batch_edge_gradient_device", self.handle, p.data_ptr(), lam.data_ptr(),
               None if explicit_dg is None else explicit_dg.data_ptr(),
               None if dg_dparam is None else dg_dparam.data_ptr(), out.data_ptr())
        return out

    def edge_flows(self, p, resistance, quad_coeff=None, flow=None):
        """Edge flows from device pressures p, as update_from_solution forms
        them (wrapper.rs:349-428): (p_i - p_j) / (R + k |q_prev|) with the
        previous flows taken from ``flow`` (required when quad_coeff is given)
        and the new ones written there.  Float64 CUDA tensors; returns flow."""
        import torch

        _require_cuda_f64(p, self.n_rows, "p")
        _require_cuda_f64(resistance, self.edge_i.size, "resistance")
        if quad_coeff is not None:
            _require_cuda_f64(quad_coeff, self.edge_i.size, "quad_coeff")
            if flow is None:
                raise N.CfdlaError(N.ERR_INVALID_ARG, "a quadratic coefficient needs the previous flows")
        if flow is None:
            flow = torch.zeros(self.edge_i.size, dtype=torch.float64, device=p.device)
        _require_cuda_f64(flow, self.edge_i.size, "flow")
        bind_torch_stream()
        N.call("cfdla_batch_edge_flows_device", self.handle, p.data_ptr(), resistance.data_ptr(),
               None if quad_coeff is None else quad_coeff.data_ptr(), flow.data_ptr())
        return flow

    def outlet_flux(self, p, conductance, outlet_mask, derivatives: bool = False):
        """Per-network flux into the nodes flagged in outlet_mask (uint8 CUDA
        tensor per node) from device pressures and conductances; with
        derivatives, also dJ/dp (per node, 0 on Dirichlet nodes) and the
        explicit dJ/dG (per edge)."""
        import torch

        _require_cuda_f64(p, self.n_rows, "p")
        _require_cuda_f64(conductance, self.edge_i.size, "conductance")
        if not (_is_torch(outlet_mask) and outlet_mask.is_cuda and outlet_mask.dtype == torch.uint8
                and outlet_mask.is_contiguous() and outlet_mask.numel() == self.n_rows):
            raise N.CfdlaError(N.ERR_INVALID_ARG, "outlet_mask must be a contiguous uint8 CUDA tensor of n_rows")
        J = torch.empty(self.n_sys, dtype=torch.float64, device=p.device)
        dp = torch.empty_like(p) if derivatives else None
        dg = torch.empty_like(conductance) if derivatives else None
        bind_torch_stream()
        N.call("cfdla_batch_outlet_flux_device", self.handle, p.data_ptr(), conductance.data_ptr(),
               outlet_mask.data_ptr(), J.data_ptr(), None if dp is None else dp.data_ptr(),
               None if dg is None else dg.data_ptr())
        return (J, dp, dg) if derivatives else J

    # ---- autotuning -------------------------------------------------------------------
    DEFAULT_CANDIDATES = [(w, b) for w in (0, 64, 128, 256, 512) for b in (1, 2, 4)]

    def signature(self, config: IterativeSolverConfig) -> str:
        sizes = np.diff(self.node_offset)
        hist = np.bincount(np.minimum(np.log2(np.maximum(sizes, 1)).astype(int), 12), minlength=13)
        return f"{int(config.precision)}|{config.tolerance:g}|" + ",".join(str(int(h)) for h in hist)

    def autotune(self, config: IterativeSolverConfig | None = None, candidates=None, cache_path: str | None = None):
        """Time candidate (wave_max_rows, blocks_per_cu) schedules with a full
        solve each, apply and return the fastest; results are cached in JSON
        keyed by precision, tolerance and a histogram of network sizes."""
        config = config or IterativeSolverConfig.new(1e-10).with_max_iterations(20000)
        cache_path = cache_path or os.path.join(os.path.expanduser("~"), ".cache", "cfdrs_amd",
                                                "batch_autotune.json")
        key = self.signature(config)
        cache = {}
        if os.path.exists(cache_path):
            with open(cache_path) as f:
                cache = json.load(f)
        if key in cache:
            self.schedule = cache[key]["best"]
            return cache[key]
        cands = candidates or self.DEFAULT_CANDIDATES
        arr = (N.BatchSchedule * len(cands))(*[N.BatchSchedule(w, b) for w, b in cands])
        ms = (C.c_double * len(cands))()
        best = N.BatchSchedule()
        cfg = config.native(N.KIND_CG)
        N.call("cfdla_batch_autotune", self.handle, C.byref(cfg), len(cands), arr, ms, C.byref(best))
        entry = {"best": {"wave_max_rows": best.wave_max_rows, "blocks_per_cu": best.blocks_per_cu},
                 "timings_ms": {f"{w}/{b}": ms[k] for k, (w, b) in enumerate(cands)}}
        cache[key] = entry
        os.makedirs(os.path.dirname(cache_path), exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        return entry

    def close(self) -> None:
        if self._h is not None and N._lib is not None:
            N.lib().cfdla_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def channel_batch(nets: SynthNetworks, inlet_pressure: float = INLET_PRESSURE, height: float = CHANNEL_HEIGHT,
                  viscosity: float = WATER_VISCOSITY) -> NetworkBatch:
    """Batch with node 0 at the inlet pressure and the last node grounded."""
    g, _ = rect_conductance(nets.width, height, nets.length, viscosity)
    n_total = int(nets.node_offset[-1])
    is_dir = np.zeros(n_total, np.uint8)
    dval = np.zeros(n_total)
    is_dir[nets.node_offset[:-1]] = 1
    dval[nets.node_offset[:-1]] = inlet_pressure
    is_dir[nets.node_offset[1:] - 1] = 1
    return NetworkBatch(nets.node_offset, nets.edge_offset, nets.edge_i, nets.edge_j, g, is_dir, dval)


def outlet_flux(nets: SynthNetworks, g, p):
    """J_s = sum over edges touching the outlet (last node) of G_e (p_other - p_outlet), per network."""
    sys_of_edge = np.repeat(np.arange(nets.count), np.diff(nets.edge_offset))
    outlet = nets.node_offset[1:] - 1 - nets.node_offset[:-1]  # local id
    o = outlet[sys_of_edge]
    base = nets.node_offset[:-1][sys_of_edge]
    at_i = nets.edge_i == o
    at_j = nets.edge_j == o
    other = np.where(at_i, nets.edge_j, nets.edge_i) + base
    touch = at_i | at_j
    contrib = np.where(touch, g * (p[other] - p[o + base]), 0.0)
    return np.bincount(sys_of_edge, weights=contrib, minlength=nets.count)


class WidthDesign:
    """Outlet flux of every network as a function of the channel widths, on
    the GPU: widths -> rectangular-duct conductances (rectangular.rs:95-157)
    -> device re-assembly -> batched Jacobi-PCG -> per-network outlet flux.
    ``flux(width)`` is a differentiable torch function whose backward pass is
    one batched adjoint solve (A is symmetric on the free nodes; Dirichlet
    rows are identity rows, so the adjoint vanishes there) and the edge
    gradient kernel, times dG/dw.  No reference counterpart (the reference
    has no adjoint code)."""

    def __init__(self, batch: NetworkBatch, height, length, outlet_mask, viscosity=WATER_VISCOSITY,
                 config: IterativeSolverConfig | None = None):
        import torch

        self.batch = batch
        self.height = height
        self.length = length
        self.viscosity = viscosity
        self.outlet_mask = outlet_mask
        self.config = config or IterativeSolverConfig.new(1e-12).with_max_iterations(50000)
        dev = outlet_mask.device
        self.sys_of_node = torch.from_numpy(np.repeat(np.arange(batch.n_sys), np.diff(batch.node_offset))).to(dev)
        self.sys_of_edge = torch.from_numpy(np.repeat(np.arange(batch.n_sys), np.diff(batch.edge_offset))).to(dev)
        self._g = None  # the conductances the batch currently holds

    def _set(self, g) -> None:
        if self._g is None or self._g.data_ptr() != g.data_ptr():
            self.batch.update_conductance_device(g)
            self._g = g

    def _solve(self, x, rhs=None) -> None:
        _, reps = self.batch.solve(x, self.config, rhs=rhs)
        bad = [r for r in reps if r.termination not in (Termination.Converged, Termination.InitialResidual)]
        if bad:
            raise RuntimeError(f"{len(bad)} network solve(s) did not converge: {bad[0]}")

    def forward(self, width):
        """(J per network, p, G, dG/dw) for a float64 CUDA tensor of widths."""
        import torch

        g, dg = rect_conductance_device(width.detach(), self.height, self.length, self.viscosity)
        self._set(g)
        p = torch.zeros(self.batch.n_rows, dtype=torch.float64, device=width.device)
        self._solve(p)
        return self.batch.outlet_flux(p, g, self.outlet_mask), p, g, dg

    def gradient(self, weights, p, g, dg):
        """d(sum_s weights_s J_s)/dw: one batched adjoint solve and the edge
        gradient kernel."""
        import torch

        self._set(g)
        _, dJdp, dJdg = self.batch.outlet_flux(p, g, self.outlet_mask, derivatives=True)
        rhs = (weights[self.sys_of_node] * dJdp).contiguous()
        lam = torch.zeros_like(p)
        self._solve(lam, rhs)
        explicit = (weights[self.sys_of_edge] * dJdg).contiguous()
        return self.batch.edge_gradient(p, lam, explicit, dg)

    def flux(self, width):
        """Differentiable per-network outlet flux J(width)."""
        import torch

        design = self

        class OutletFluxOfWidths(torch.autograd.Function):
            @staticmethod
            def forward(ctx, w):
                _require_cuda_f64(w, design.batch.edge_i.size, "width")
                J, p, g, dg = design.forward(w)
                ctx.save_for_backward(p, g, dg)
                return J

            @staticmethod
            def backward(ctx, grad_out):
                p, g, dg = ctx.saved_tensors
                return design.gradient(grad_out.contiguous(), p, g, dg)

        return OutletFluxOfWidths.apply(width)
