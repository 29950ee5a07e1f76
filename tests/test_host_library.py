"""CPU tests of libcfdla: the C ABI exports every symbol include/cfdla.h
declares, host-side assembly and AMG setup are bit-exact against the oracle,
and the host API mirrors the reference's error and outcome semantics.
No device calls."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import cfdrs_amd as ca
from cfdrs_amd import _native as N
from cfdrs_amd import problems as P
from cfdrs_amd.plugin import PluginRegistry, capabilities, register_backend
from conftest import ROOT, grid_laplacian, dense_to_csr
from oracle import oracle as O


def header_symbols():
    text = open(os.path.join(ROOT, "include", "cfdla.h")).read()
    return sorted(set(re.findall(r"\b(cfdla_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = N.lib()
    syms = header_symbols()
    assert len(syms) >= 29
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # every declared symbol has a ctypes signature in the binding
    assert not [s for s in syms if s not in N.SIGNATURES], [s for s in syms if s not in N.SIGNATURES]


def test_version_and_error_string():
    assert b"gfx950" in N.lib().cfdla_version()
    with pytest.raises(N.CfdlaError) as e:
        ca.CsrMatrix.from_parts([1.0], [5], [0, 1], 1, 1)  # column out of range
    assert e.value.status == N.ERR_INVALID_ARG
    assert "column index out of range" in str(e.value)


def test_csr_validation():
    with pytest.raises(N.CfdlaError):
        ca.CsrMatrix.from_parts([1.0, 2.0], [0, 0], [0, 2], 1, 1)  # duplicate
    with pytest.raises(N.CfdlaError):
        ca.CsrMatrix.from_parts([1.0], [0], [0, 2], 1, 1)  # row_ptr beyond nnz
    m = ca.CsrMatrix.from_parts([3.0, 1.0], [1, 0], [0, 2], 1, 2)  # unsorted row is accepted and sorted
    assert m.col_indices.tolist() == [0, 1] and m.values.tolist() == [1.0, 3.0]


def test_empty_matrix_accepted():
    m = ca.CsrMatrix.from_parts([], [], [0], 0, 0)
    assert m.nnz == 0 and m.shape == (0, 0)


def test_synth_uniform_reproducible_and_in_range():
    a = P.uniform(42, 10000)
    assert np.array_equal(a, P.uniform(42, 10000))
    assert a.min() >= -1.0 and a.max() < 1.0 and abs(a.mean()) < 0.05
    assert not np.array_equal(a, P.uniform(43, 10000))


def test_momentum_assembly_matches_oracle():
    n = 40
    u, v = P.cavity_velocity(n, seed=7)
    h = 1.0 / (n - 1)
    rp, ci, vals = P.momentum_u(n, h, h, 0.05, 1e-3, 1.0, u, v)
    o = O.assemble_momentum_u(n, n, h, h, 0.05, 1e-3, 1.0, u, v).arrays()
    assert np.array_equal(rp, o[0]) and np.array_equal(ci, o[1]) and np.array_equal(vals, o[2])


def test_pressure_rhs_vanishes_for_divergence_free_field():
    # u = d psi/dy, v = -d psi/dx with psi = x^2 y + 3 x y^2: central differences are exact
    n = 30
    h = 1.0 / (n - 1)
    x, y = np.meshgrid(np.linspace(0, 1, n), np.linspace(0, 1, n))  # [j, i]
    u = (x ** 2 + 6 * x * y).ravel()
    v = -(2 * x * y + 3 * y ** 2).ravel()
    assert np.max(np.abs(P.pressure_rhs(n, h, h, 1.0, 0.1, u, v))) < 1e-10
    assert P.pressure_rhs(n, h, h, 1.0, 0.1, u, v, pin=False).size == (n - 2) ** 2


@pytest.mark.parametrize("n,dirichlet", [(16, (0, 1, 0, 0)), (7, (1, 0, 1, 1)), (1, (0, 0, 0, 0)),
                                         (9, (0, 0, 0, 0))])
def test_pressure_correction_assembly_matches_oracle(n, dirichlet):
    h = 1.0 / n
    rp, ci, v = P.pressure_correction_2d(n, n + 3, h, h * 0.5, dirichlet)
    o = O.assemble_pressure_correction(n, n + 3, h, h * 0.5, dirichlet).arrays()
    assert np.array_equal(rp, o[0]) and np.array_equal(ci, o[1]) and np.array_equal(v, o[2])


def test_config1_shape():
    rp, ci, v, b = P.config1(128)
    assert rp.size == 16385 and ci.size == 81408 and b.size == 16384


@pytest.mark.parametrize("dims,mask", [((8, 8, 8), 1 << 4), ((9, 6, 1), 0b0011), ((4, 5, 3), 0b111111)])
def test_varcoef_assembly_matches_oracle(dims, mask):
    nx, ny, nz = dims
    k = np.exp(P.uniform(99, nx * ny * nz, -2.3, 2.3))
    dv = np.array([1.0, 0.0, 0.5, -0.5, 2.0, 0.0])
    rp, ci, v, bc = P.varcoef(nx, ny, nz, 7.0, k, mask, dv)
    a, rhs = O.assemble_varcoef(nx, ny, nz, 7.0, k, mask, dv)
    o = a.arrays()
    assert np.array_equal(rp, o[0]) and np.array_equal(ci, o[1]) and np.array_equal(v, o[2])
    assert np.array_equal(bc, rhs)


def _compare_hierarchy(amg, oamg):
    st = amg.statistics
    assert st.num_levels == oamg.num_levels
    for l in range(st.num_levels):
        for w in ("A", "R", "P") if l + 1 < st.num_levels else ("A",):
            a = amg.level(l, w)
            o = oamg.level(l, w).arrays()
            assert np.array_equal(a[0], o[0]), (l, w, "row_ptr")
            assert np.array_equal(a[1], o[1]), (l, w, "col")
            assert np.array_equal(a[2], o[2]), (l, w, "values")
        if l + 1 < st.num_levels:
            assert np.array_equal(amg.level_map(l), oamg.level_map(l))
    op, gr = oamg.complexity()
    assert st.operator_complexity == op and st.grid_complexity == gr


@pytest.mark.parametrize("interp", ["direct", "smoothed"])
def test_amg_aggregation_hierarchy_bit_exact(interp):
    n = 12
    k = P.config2_k(n)
    rp, ci, v, _ = P.varcoef(n, n, n, float(n * n), k, 1 << 4, np.zeros(6))
    m = ca.CsrMatrix.from_parts(v, ci, rp, n ** 3, n ** 3)
    cfg = ca.AMGConfig.gpu_aggregation(smoothed=interp == "smoothed")
    amg = ca.AlgebraicMultigrid(m, cfg)
    oamg = O.Amg(O.Csr.from_arrays(rp, ci, v, n ** 3), coarsening="aggregation", interp=interp, smoother="jacobi",
                 coarse_solve="direct", sa_omega=cfg.smoothing_omega)
    _compare_hierarchy(amg, oamg)


@pytest.mark.parametrize("coarsening", ["rs", "hybrid"])
def test_amg_classical_hierarchy_bit_exact(coarsening):
    rp, ci, v, _ = P.config1(20)
    m = ca.CsrMatrix.from_parts(v, ci, rp, 400, 400)
    cfg = ca.AMGConfig(smoother_type=ca.SmootherType.Jacobi,
                       coarsening_strategy={"rs": ca.CoarseningStrategy.RugeStueben,
                                            "hybrid": ca.CoarseningStrategy.Hybrid}[coarsening])
    amg = ca.AlgebraicMultigrid(m, cfg)
    oamg = O.Amg(O.Csr.from_arrays(rp, ci, v, 400), coarsening=coarsening, smoother="jacobi")
    _compare_hierarchy(amg, oamg)


def test_amg_recompute_keeps_transfer_operators():
    # amg.rs:179-214: new values, same pattern; P and R unchanged, A_c recomputed
    n = 10
    rp, ci, v, _ = P.varcoef(n, n, 1, 100.0, np.exp(P.uniform(5, n * n)), 0b11, np.zeros(6))
    m = ca.CsrMatrix.from_parts(v, ci, rp, n * n, n * n)
    cfg = ca.AMGConfig.gpu_aggregation(smoothed=False)
    cfg.min_coarse_size = 10
    amg = ca.AlgebraicMultigrid(m, cfg)
    p0 = amg.level(0, "P")
    v2 = v * 1.5
    m2 = ca.CsrMatrix.from_parts(v2, ci, rp, n * n, n * n)
    amg.recompute(m2)
    p1 = amg.level(0, "P")
    assert all(np.array_equal(a, b) for a, b in zip(p0[:3], p1[:3]))
    oamg = O.Amg(O.Csr.from_arrays(rp, ci, v2, n * n), coarsening="aggregation", interp="direct",
                 smoother="jacobi", coarse_solve="direct", min_coarse=10)
    # values of the same transfer operators: identical Galerkin products
    for l in range(1, amg.statistics.num_levels):
        a, o = amg.level(l, "A"), oamg.level(l, "A").arrays()
        assert np.array_equal(a[2], o[2])


def test_amg_default_config_builds_and_bad_smoothers_are_rejected():
    rp, ci, v, _ = P.config1(8)
    m = ca.CsrMatrix.from_parts(v, ci, rp, 64, 64)
    amg = ca.AlgebraicMultigrid(m, ca.AMGConfig())  # the reference default: RS, classical, Gauss-Seidel
    assert amg.statistics.num_levels >= 1
    cfg = ca.AMGConfig()
    cfg.relaxation_factor = 0.0
    with pytest.raises(N.CfdlaError) as e:
        ca.AlgebraicMultigrid(m, cfg)
    assert e.value.status == N.ERR_CONFIG
    native = ca.AMGConfig().native()
    native.smoother = 7
    import ctypes as C

    h = C.c_void_p()
    assert N.lib().cfdla_precond_create_amg(m.handle, C.byref(native), C.byref(h)) == N.ERR_CONFIG


def test_sparse_matrix_builder_sums_duplicates_in_order():
    b = ca.SparseMatrixBuilder(2, 3)
    b.add_entry(1, 2, 1.0)
    b.add_entry(0, 1, 2.0)
    b.add_entry(1, 2, 0.5)
    b.add_entry(1, 0, -1.0)
    rp, ci, v = b.build_arrays()
    assert rp.tolist() == [0, 1, 3] and ci.tolist() == [1, 0, 2] and v.tolist() == [2.0, -1.0, 1.5]
    with pytest.raises(N.CfdlaError):
        b.add_entry(2, 0, 1.0)


def test_restart_width_ladder(known):
    for req, w in known["restart_ladder"]["cases"]:
        assert ca.restart_width(req) == w
    assert ca.KrylovWorkspace(30, 10).width == 32


def test_interpret_classification():
    T = ca.Termination
    mk = lambda t: ca.SolveReport(3, 1.0, 0.1, 0.5, t)
    assert ca.interpret("x", mk(T.Converged)).converged()
    assert ca.interpret("x", mk(T.InitialResidual)).converged()
    assert ca.interpret("x", mk(T.MaxIterations)).kind == ca.SolveOutcome.Stalled
    assert ca.interpret("x", mk(T.Breakdown)).kind == ca.SolveOutcome.BrokenDown
    assert ca.interpret("x", mk(T.NonPositiveCurvature)).kind == ca.SolveOutcome.BrokenDown
    with pytest.raises(ca.SolverError):
        ca.interpret("x", ca.SolveError(N.CfdlaError(N.ERR_DIMENSION, "bad")))
    assert ca.converged_or_none("x", mk(T.MaxIterations)) is None
    assert ca.converged_or_none("x", mk(T.Converged)) is not None


def test_solver_kind_names():
    assert ca.SolverKind.ConjugateGradient().name() == "ConjugateGradient"
    assert ca.SolverKind.BiCgStab().name() == "BiCGSTAB"
    assert ca.SolverKind.Gmres(30).name() == "GMRES"
    assert ca.SolverKind.Gmres(30) == ca.SolverKind.Gmres(30) != ca.SolverKind.Gmres(8)


def test_plugin_surface():
    reg = PluginRegistry()
    p = register_backend("hip", reg)
    assert reg.get("hip-linear-solver") is p and "amg" in capabilities()
    with pytest.raises(ValueError):
        register_backend("wgpu")
    with pytest.raises(ValueError):
        reg.register(p)


def test_amg_recompute_with_unsorted_matrix_passes_creation_order():
    # a matrix created with unsorted rows: recompute hands the library the
    # values in creation order (a fresh array that must stay referenced
    # across the call); level 0 then holds the new values in stored order
    rng = np.random.default_rng(3)
    n = 400
    rows = np.repeat(np.arange(n), 5)
    cols = np.concatenate([rng.choice(n, 5, replace=False) for _ in range(n)])
    vals = -rng.uniform(0.1, 1.0, rows.size)
    diag = np.zeros(n)
    np.add.at(diag, rows, -vals)
    keep = rows != cols
    rows = np.concatenate([rows[keep], np.arange(n)])
    cols = np.concatenate([cols[keep], np.arange(n)])
    vals = np.concatenate([vals[keep], diag + 1.0])
    order = np.lexsort((rng.random(rows.size), rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    rp = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n))])
    a = ca.CsrMatrix.from_parts(vals, cols, rp, n, n)
    assert a._order is not None  # the rows really were unsorted
    cfg = ca.AMGConfig.smoothed_aggregation()
    cfg.min_coarse_size = 20
    amg = ca.AlgebraicMultigrid(a, cfg)
    L = amg.statistics.num_levels
    coarse0 = amg.level(L - 1, "A")[2].copy()
    a.update_values(vals * 1.5)
    amg.recompute(a)
    assert np.array_equal(amg.level(0, "A")[2], a.values)
    assert np.allclose(amg.level(L - 1, "A")[2], 1.5 * coarse0, rtol=1e-13, atol=0)


def test_solve_entry_points_reject_mismatched_preconditioner():
    # cfdla_solve_device and cfdla_solve_transpose_device check the preconditioner's size before
    # any device work: an AMG built for 16 rows against a 64-row matrix is a dimension error
    rp, ci, v, _ = P.config1(8)
    m = ca.CsrMatrix.from_parts(v, ci, rp, 64, 64)
    rp2, ci2, v2, _ = P.config1(4)
    cfg = ca.AMGConfig.smoothed_aggregation()
    cfg.min_coarse_size = 4
    amg = ca.AlgebraicMultigrid(ca.CsrMatrix.from_parts(v2, ci2, rp2, 16, 16), cfg)
    sc = ca.IterativeSolverConfig.new(1e-10).native(N.KIND_CG)
    rep = N.Report()
    for fn in ("cfdla_solve_device", "cfdla_solve_transpose_device"):
        st = getattr(N.lib(), fn)(m.handle, amg.handle, None, None, C.byref(sc), C.byref(rep))
        assert st == N.ERR_DIMENSION, fn
        assert b"preconditioner size differs" in N.lib().cfdla_last_error()


class _FakeTensor:
    """Quacks like a torch tensor for the boundary checks; any native call
    through data_ptr() would fail the test."""

    def __init__(self, n, is_cuda=True, contiguous=True, dtype=None):
        import torch

        self.n, self.is_cuda, self._contig = n, is_cuda, contiguous
        self.dtype = dtype or torch.float64
        self.device = "cuda:0" if is_cuda else "cpu"

    def numel(self):
        return self.n

    def is_contiguous(self):
        return self._contig

    def data_ptr(self):
        raise AssertionError("a native call was about to be made")


@pytest.mark.parametrize("bad", ["cpu", "strided"])
def test_device_entry_points_reject_cpu_and_strided_tensors(bad):
    # a CPU tensor or a strided view must never reach cfdla_solve_device, cfdla_spmv_device or
    # cfdla_precond_apply_device: the kernels would read host memory or a dense stride
    rp, ci, v, _ = P.config1(8)
    m = ca.CsrMatrix.from_parts(v, ci, rp, 64, 64)
    cfg = ca.AMGConfig.smoothed_aggregation()
    cfg.min_coarse_size = 4
    amg = ca.AlgebraicMultigrid(m, cfg)
    good = _FakeTensor(64)
    wrong = _FakeTensor(64, is_cuda=bad != "cpu", contiguous=bad != "strided")
    conf = ca.IterativeSolverConfig.new(1e-10)
    for args in ((wrong, good), (good, wrong)):
        with pytest.raises(ca.SolveError) as e:
            ca.cg_preconditioned(m, args[0], amg, args[1], conf)
        assert e.value.status == N.ERR_INVALID_ARG
        with pytest.raises(N.CfdlaError) as e:
            m.spmv_into(*args)
        assert e.value.status == N.ERR_INVALID_ARG
        with pytest.raises(N.CfdlaError) as e:
            amg.apply_to(*args)
        assert e.value.status == N.ERR_INVALID_ARG
