"""GPU edge cases of the C ABI, against the oracle or closed forms: empty
(0 x 0) systems, 1 x 1 systems through every solver and precision, rows with
no entries, a non-finite right-hand side, the 4095 / 4096-row boundary where
stencil matrices switch to the diagonal kernel, and an AMG hierarchy of a
single level.  Integer data keeps products exact (bit-exact bar); solves meet
an explicit FP64 relative residual of 1e-10."""
import numpy as np
import pytest

import cfdrs_amd as ca
from oracle import oracle as O

pytestmark = pytest.mark.gpu

KINDS = ["cg", "bicgstab", "gmres"]
PRECISIONS = [ca.Precision.FP64, ca.Precision.FP32_REFINED]


def _solve(kind, a, b, prec, x, cfg):
    if kind == "cg":
        return ca.cg_preconditioned(a, b, prec, x, cfg)
    if kind == "bicgstab":
        return ca.bicgstab_preconditioned(a, b, prec, x, cfg)
    return ca.gmres_preconditioned(a, b, prec, x, cfg, 30)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("precision", PRECISIONS)
def test_empty_system(kind, precision):
    a = ca.CsrMatrix.from_parts([], [], [0], 0, 0)
    assert a.spmv(np.zeros(0)).shape == (0,)
    x = np.zeros(0)
    rep = _solve(kind, a, np.zeros(0), None, x, ca.IterativeSolverConfig.new(1e-10).with_precision(precision))
    assert rep.termination == ca.Termination.InitialResidual and rep.iterations == 0
    with pytest.raises(ca.CfdlaError):
        ca.AlgebraicMultigrid(a, ca.AMGConfig.smoothed_aggregation())


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("prec", ["identity", "jacobi", "amg"])
def test_one_by_one(kind, precision, prec):
    a = ca.CsrMatrix.from_parts([4.0], [0], [0, 1], 1, 1)
    m = {"identity": None, "jacobi": ca.JacobiPreconditioner(a),
         "amg": ca.AlgebraicMultigrid(a, ca.AMGConfig.smoothed_aggregation())}[prec]
    x = np.zeros(1)
    rep = _solve(kind, a, np.array([2.0]), m, x, ca.IterativeSolverConfig.new(1e-10).with_precision(precision))
    assert rep.termination == ca.Termination.Converged, rep
    assert abs(x[0] - 0.5) <= 1e-10 * 0.5


def test_rows_without_entries_and_nonfinite_x():
    # a square matrix whose nonzeros sit in every third row: the other rows
    # give exact zeros whatever x holds in columns they do not reference
    n = 6000
    rows = np.arange(0, n, 3)
    rp = np.zeros(n + 1, np.int64)
    rp[rows + 1] = 1
    rp = np.cumsum(rp)
    ci = rows.astype(np.int64)
    v = np.full(rows.size, 2.0)
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    x = np.arange(n, dtype=float)
    x[1] = np.inf  # column 1 is referenced by no row
    y = a.spmv(x)
    assert np.array_equal(y, O.Csr.from_arrays(rp, ci, v, n).spmv(x))
    empty = ca.CsrMatrix.from_parts([], [], np.zeros(n + 1, np.int64), n, n)
    assert np.array_equal(empty.spmv(np.full(n, np.nan)), np.zeros(n))


def test_nonfinite_right_hand_side_is_an_error():
    a = ca.CsrMatrix.from_parts([2.0, 3.0], [0, 1], [0, 1, 2], 2, 2)
    for kind in KINDS:
        with pytest.raises(ca.SolveError):
            _solve(kind, a, np.array([1.0, np.nan]), None, np.zeros(2), ca.IterativeSolverConfig.new(1e-10))


@pytest.mark.parametrize("n", [4095, 4096, 4097])
def test_tridiagonal_at_the_diagonal_kernel_threshold(n):
    # integer 1D operator; from 4096 rows it takes the symmetric diagonal
    # kernel (2 stored diagonals), below it the row-block kernel: same values
    main = np.full(n, 4.0)
    off = -np.ones(n - 1)
    rp = np.zeros(n + 1, np.int64)
    ci, v = [], []
    for i in range(n):
        for j, w in ((i - 1, -1.0), (i, 4.0), (i + 1, -1.0)):
            if 0 <= j < n:
                ci.append(j)
                v.append(w)
        rp[i + 1] = len(ci)
    ci, v = np.array(ci, np.int64), np.array(v)
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    st = a.storage()
    if n >= 4096:
        assert st == {"format": "dia", "stored_entries": 2 * n, "units": 2}
    else:
        assert st["format"] == "csr_rowblock"
    x = np.round(np.random.default_rng(n).uniform(-50, 50, n))
    assert np.array_equal(a.spmv(x), O.Csr.from_arrays(rp, ci, v, n).spmv(x))
    xs = np.zeros(n)
    b = a.spmv(np.ones(n))
    rep = ca.cg_preconditioned(a, b, ca.JacobiPreconditioner(a), xs,
                               ca.IterativeSolverConfig.new(1e-12).with_max_iterations(500))
    assert rep.termination == ca.Termination.Converged
    assert np.max(np.abs(xs - 1.0)) <= 1e-10


def test_single_level_amg_is_the_direct_solve():
    # fewer rows than min_coarse_size: the hierarchy is the coarsest level,
    # one V-cycle is the FP64 dense solve and CG converges in one step
    rng = np.random.default_rng(4)
    m = rng.normal(size=(20, 20))
    d = m @ m.T + 20 * np.eye(20)
    rp = np.arange(0, 401, 20, dtype=np.int64)
    ci = np.tile(np.arange(20, dtype=np.int64), 20)
    a = ca.CsrMatrix.from_parts(d.ravel(), ci, rp, 20, 20)
    amg = ca.AlgebraicMultigrid(a, ca.AMGConfig.smoothed_aggregation())
    assert amg.statistics.level_sizes == [20]
    b = rng.normal(size=20)
    z = np.empty(20)
    amg.apply_to(b, z)
    assert np.allclose(z, np.linalg.solve(d, b), rtol=1e-12, atol=1e-14)
    x = np.zeros(20)
    rep = ca.cg_preconditioned(a, b, amg, x, ca.IterativeSolverConfig.new(1e-10))
    assert rep.termination == ca.Termination.Converged and rep.iterations <= 2
