"""GPU parity of the row-block CSR product in every gather layout.

quad: every thread gathers x for 4 consecutive stored entries of its block as
it loads them; rows: the block's columns and values are staged in LDS and
each row's lanes gather x while walking the row; strided: staged as for rows,
entries dealt to threads round robin, products summed per row as for quad;
sorted: the quad layout over blocks stored in column order, each product
placed back at its row-ordered position, so its sums are the quad layout's
bit for bit.  Tolerances:

* integer-valued products: bit-exact in FP64 and FP32 (every product and
  partial sum is an exact integer below 2^24);
* a stored zero times a non-finite x gives NaN, as the oracle's CSR product;
* V-cycles: 1e-11 against the oracle (FP64), 1e-5 (FP32), and the layouts
  within 1e-13 of each other (they differ by rounding only);
* solves: explicit FP64 relative residual <= 1e-10.
"""
import numpy as np
import pytest

import cfdrs_amd as ca
from cfdrs_amd import problems as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LAYOUTS = ["quad", "rows", "strided", "sorted"]


def _rows(lens, m, seed):
    rng = np.random.default_rng(seed)
    rp, ci, v = [0], [], []
    for L in lens:
        cols = np.sort(rng.choice(m, size=min(int(L), m), replace=False))
        ci.extend(cols.tolist())
        v.extend(rng.integers(-9, 10, size=cols.size).astype(float).tolist())
        rp.append(len(ci))
    return np.array(rp, np.int64), np.array(ci, np.int64), np.array(v)


def _stencil7(nn, seed):
    # 7-point pattern, integer values (diagonal 6, off-diagonals in
    # {-3..-1, 1..3}: no stored zero, so the matrix can take diagonal storage)
    rng = np.random.default_rng(seed)
    n = nn ** 3
    rp, ci, v = [0], [], []
    for k in range(nn):
        for j in range(nn):
            for i in range(nn):
                r = (k * nn + j) * nn + i
                for c, ok in ((r - nn * nn, k > 0), (r - nn, j > 0), (r - 1, i > 0), (r, True),
                              (r + 1, i < nn - 1), (r + nn, j < nn - 1), (r + nn * nn, k < nn - 1)):
                    if ok:
                        ci.append(c)
                        v.append(6.0 if c == r else float(rng.choice([-3, -2, -1, 1, 2, 3])))
                rp.append(len(ci))
    return np.array(rp, np.int64), np.array(ci, np.int64), np.array(v), n


def _case(name):
    """(row_ptr, cols, values, nrows, ncols) with integer values."""
    rng = np.random.default_rng({"stencil7": 11, "short": 12, "long": 13, "mixed": 14}[name])
    if name == "stencil7":
        rp, ci, v, n = _stencil7(20, 3)
        return rp, ci, v, n, n
    if name == "short":  # prolongator-like: 0..8 entries, every 13th row empty
        lens = rng.integers(0, 9, size=20000)
        lens[::13] = 0
        return (*_rows(lens, 30000, 5), 20000, 30000)
    if name == "long":  # 30..700 entries: 2..64 lanes per row
        lens = rng.integers(30, 701, size=700)
        return (*_rows(lens, 90001, 6), 700, 90001)
    if name == "mixed":  # rows beyond the 2048-entry cap among short ones
        lens = np.concatenate([rng.integers(0, 12, size=3000), [2049, 7, 5000, 3, 1, 4097], rng.integers(0, 40, 500)])
        return (*_rows(lens, 9000, 7), lens.size, 9000)
    raise ValueError(name)


@pytest.mark.parametrize("storage", ["csr_rowblock", "csr_int32"])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("case", ["stencil7", "short", "long", "mixed"])
def test_spmv_exact_in_every_layout(case, layout, dtype, storage):
    import torch

    rp, ci, v, n, m = _case(case)
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, m)
    assert a.set_storage(storage)["format"] == "csr_rowblock"
    assert a.set_gather(layout) == {"fp64": layout, "fp32": layout}
    x = np.random.default_rng(2).integers(-40, 41, size=m).astype(float)
    yo = O.Csr.from_arrays(rp, ci, v, n, m).spmv(x)
    dt = torch.float64 if dtype == "f64" else torch.float32
    y = a.spmv(torch.tensor(x, dtype=dt, device="cuda")).double().cpu().numpy()
    assert np.array_equal(y, yo)


@pytest.mark.parametrize("layout", LAYOUTS)
def test_spmv_explicit_zero_with_non_finite_x(layout):
    nn = 48
    n = nn * nn
    rp, ci, v, _ = P.varcoef(nn, nn, 1, 1.0, np.ones(n), 0b0001, np.zeros(6))
    v = v.copy()
    v[rp[nn + 5] + 1] = 0.0
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    a.set_storage("csr_rowblock")
    a.set_gather(layout)
    x = np.round(P.uniform(7, n, -30.0, 30.0))
    x[ci[rp[nn + 5] + 1]] = np.inf
    orc = O.Csr.from_arrays(rp, ci, v, n)
    y, yo = a.spmv(x), orc.spmv(x)
    assert np.isnan(y[nn + 5])
    assert np.array_equal(np.isnan(y), np.isnan(yo))
    fin = np.isfinite(yo)
    assert np.array_equal(y[fin], yo[fin])


def test_gather_auto_choice():
    # the 7-point operator gathers for consecutive rows (its rows' k-th
    # entries are consecutive columns); diagonal storage reports quad
    rp, ci, v, n = _stencil7(24, 4)
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    assert a.storage()["format"] == "dia"
    assert a.gather() == {"fp64": "quad", "fp32": "quad"}
    a.set_storage("csr_int32")
    assert a.gather() == {"fp64": "rows", "fp32": "rows"}
    a.set_gather("quad")
    assert a.gather() == {"fp64": "quad", "fp32": "quad"}
    a.set_gather("auto")
    assert a.gather() == {"fp64": "rows", "fp32": "rows"}
    assert a.set_gather("strided") == {"fp64": "strided", "fp32": "strided"}
    assert a.set_gather("sorted") == {"fp64": "sorted", "fp32": "sorted"}
    with pytest.raises(ValueError):
        a.set_gather("columns")


@pytest.mark.parametrize("nn", [20, 23])
def test_amg_vcycle_in_every_layout(nn):
    # every row-block product of the hierarchy (restrictions, prolongations,
    # row-block levels, the fused coarse product) in each layout
    import torch

    n = nn ** 3
    rp, ci, v, _ = P.varcoef(nn, nn, nn, float(nn * nn), P.config2_k(nn), 1 << 4, np.zeros(6))
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    cfg = ca.AMGConfig.smoothed_aggregation()
    cfg.min_coarse_size = 100
    oamg = O.Amg(O.Csr.from_arrays(rp, ci, v, n), coarsening="three_phase", interp="smoothed", smoother="jacobi",
                 omega=cfg.relaxation_factor, coarse_solve="direct", sa_omega=cfg.smoothing_omega,
                 theta=cfg.strength_threshold, min_coarse=cfg.min_coarse_size)
    r = P.uniform(79, n)
    zo = oamg.apply(r)
    rt = torch.tensor(r, dtype=torch.float32, device="cuda")
    out = {}
    for layout in LAYOUTS + ["auto"]:
        amg = ca.AlgebraicMultigrid(a, cfg)
        amg.set_gather(layout)
        z = np.empty(n)
        amg.apply_to(r, z)
        assert np.linalg.norm(z - zo) <= 1e-11 * np.linalg.norm(zo), layout
        z32 = torch.empty_like(rt)
        amg.apply_to(rt, z32)
        z32 = z32.double().cpu().numpy()
        assert np.linalg.norm(z32 - zo) <= 1e-5 * np.linalg.norm(zo), layout
        # a layout change after the device copies exist rebuilds them
        amg.set_gather("quad" if layout == "rows" else "rows")
        z2 = np.empty(n)
        amg.apply_to(r, z2)
        assert np.linalg.norm(z2 - zo) <= 1e-11 * np.linalg.norm(zo), layout
        out[layout] = z
    for layout in ("rows", "strided"):
        assert np.linalg.norm(out["quad"] - out[layout]) <= 1e-13 * np.linalg.norm(out["quad"])
    # column-sorted blocks sum their products exactly as the quad layout does
    assert np.array_equal(out["quad"], out["sorted"])


@pytest.mark.parametrize("case", ["short", "mixed", "stencil7"])
def test_sorted_layout_matches_quad_bit_for_bit(case):
    # real values: the column-sorted blocks must place every product where
    # the quad layout sums it, so the two agree in every bit
    import torch

    rp, ci, v, n, m = _case(case)
    v = v + np.random.default_rng(4).uniform(-0.5, 0.5, size=v.size)
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, m)
    a.set_storage("csr_rowblock")
    x = np.random.default_rng(5).standard_normal(m)
    for dt in (torch.float64, torch.float32):
        xt = torch.tensor(x, dtype=dt, device="cuda")
        a.set_gather("quad")
        yq = a.spmv(xt).cpu().numpy()
        assert a.set_gather("sorted") == {"fp64": "sorted", "fp32": "sorted"}
        ys = a.spmv(xt).cpu().numpy()
        assert np.array_equal(yq, ys), dt


@pytest.mark.parametrize("ncols", [1 << 21, (1 << 21) + 1])
def test_sorted_layout_column_limit(ncols):
    # columns are packed in 21 bits: 2^21 columns take the sorted layout,
    # one more falls back to quad; both stay exact
    rng = np.random.default_rng(10)
    lens = rng.integers(0, 40, size=300)
    rp, ci, v = _rows(lens, ncols, 11)
    ci[rp[5]] = ncols - 1 if lens[5] > 0 else ci[rp[5]]
    ci = np.concatenate([ci, [0, ncols - 1]])
    v = np.concatenate([v, [3.0, -2.0]])
    rp = np.concatenate([rp, [rp[-1] + 2]])
    n = rp.size - 1
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, ncols)
    a.set_storage("csr_rowblock")
    want = "sorted" if ncols <= (1 << 21) else "quad"
    assert a.set_gather("sorted") == {"fp64": want, "fp32": want}
    x = rng.integers(-20, 21, size=ncols).astype(float)
    assert np.array_equal(a.spmv(x), O.Csr.from_arrays(rp, ci, v, n, ncols).spmv(x))


def test_sorted_layout_device_value_refresh():
    # a device value refresh rewrites the column-sorted copies through their
    # slot map; integer values keep the products exact
    import torch

    rp, ci, v, n, m = _case("mixed")
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, m)
    a.set_storage("csr_rowblock")
    a.set_gather("sorted")
    x = np.random.default_rng(8).integers(-20, 21, size=m).astype(float)
    assert np.array_equal(a.spmv(x), O.Csr.from_arrays(rp, ci, v, n, m).spmv(x))
    v2 = np.random.default_rng(9).integers(-9, 10, size=v.size).astype(float)
    a.update_values_device(torch.tensor(v2, dtype=torch.float64, device="cuda"))
    yo = O.Csr.from_arrays(rp, ci, v2, n, m).spmv(x)
    assert np.array_equal(a.spmv(x), yo)
    y32 = a.spmv(torch.tensor(x, dtype=torch.float32, device="cuda")).double().cpu().numpy()
    assert np.array_equal(y32, yo)


@pytest.mark.parametrize("precision", [ca.Precision.FP64, ca.Precision.FP32_REFINED])
@pytest.mark.parametrize("layout", LAYOUTS)
def test_krylov_solves_in_every_layout(layout, precision):
    # 2D variable-coefficient operator, Dirichlet west and east walls
    nn = 96
    n = nn * nn
    rp, ci, v, _ = P.varcoef(nn, nn, 1, float(nn * nn), np.exp(P.uniform(5, n, -1.0, 1.0)), 0b0011, np.zeros(6))
    b = P.uniform(6, n)
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    a.set_storage("csr_rowblock")
    a.set_gather(layout)
    orc = O.Csr.from_arrays(rp, ci, v, n)
    cfg = ca.IterativeSolverConfig.new(1e-10).with_max_iterations(4000).with_precision(precision)
    for solve in (lambda x: ca.cg_preconditioned(a, b, ca.JacobiPreconditioner(a), x, cfg),
                  lambda x: ca.bicgstab_preconditioned(a, b, ca.JacobiPreconditioner(a), x, cfg),
                  lambda x: ca.gmres_preconditioned(a, b, ca.JacobiPreconditioner(a), x, cfg, 30)):
        x = np.zeros(n)
        rep = solve(x)
        assert rep.termination == ca.Termination.Converged, rep
        assert np.linalg.norm(b - orc.spmv(x)) <= 1e-10 * np.linalg.norm(b)
