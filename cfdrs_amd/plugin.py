"""Plugin surface for the GPU linear-solve backend.

Mirrors ``cfd_core::management::plugin::Plugin`` (crates/cfd-core/src/management/plugin/traits.rs:6-55)
and the registry's ``register`` / lookup (registry.rs:33-60).  Backend
selection mirrors ``ComputeDispatcher::execute`` (crates/cfd-core/src/compute/dispatch.rs:51-72):
requesting an unavailable backend is an error, never a silent fallback.
"""
from __future__ import annotations

from . import _native as N


class HipLinearSolverPlugin:
    """The MI355X sparse linear-solve backend as a CFDrs plugin."""

    def name(self) -> str:
        return "hip-linear-solver"

    def version(self) -> str:
        return N.lib().cfdla_version().decode()

    def description(self) -> str:
        return ("CSR SpMV, PCG / BiCGSTAB / GMRES(m), Jacobi and aggregation-AMG preconditioning, "
                "FP32-inner / FP64-refined solves, adjoint gradients and batched network solves on gfx950")

    def author(self):
        return None

    def license(self):
        return None

    def dependencies(self) -> list:
        return []

    def capabilities(self) -> list:
        return ["csr_spmv", "cg", "bicgstab", "gmres", "jacobi", "amg", "fp32_refined", "adjoint",
                "batched_networks"]

    def supports_parallel(self) -> bool:
        return True

    def supports_adaptive_timestep(self) -> bool:
        return False

    def supports_multiphase(self) -> bool:
        return False

    def supports_turbulence(self) -> bool:
        return False


class PluginRegistry:
    def __init__(self):
        self._plugins: dict = {}

    def register(self, plugin) -> None:
        name = plugin.name()
        if name in self._plugins:
            raise ValueError(f"Plugin '{name}' is already registered")
        for dep in plugin.dependencies():
            if dep not in self._plugins:
                raise ValueError(f"Plugin '{name}' depends on unregistered plugin '{dep}'")
        self._plugins[name] = plugin

    def get(self, name: str):
        return self._plugins.get(name)

    def list_plugins(self) -> list:
        return sorted(self._plugins)


_BACKENDS = {"hip": HipLinearSolverPlugin}


def register_backend(name: str = "hip", registry: PluginRegistry | None = None):
    """Select a compute backend by name; only "hip" exists and it must load."""
    if name not in _BACKENDS:
        raise ValueError(f"unknown backend '{name}' (available: {sorted(_BACKENDS)})")
    N.lib()  # raises NativeLibraryMissing: no fallback backend
    plugin = _BACKENDS[name]()
    if registry is not None:
        registry.register(plugin)
    return plugin


def capabilities() -> list:
    return HipLinearSolverPlugin().capabilities()
