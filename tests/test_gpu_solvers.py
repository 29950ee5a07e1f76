"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs.  Tolerances are stated per test:

* integer-valued SpMV: bit-exact (every partial sum is an exact integer);
* real-valued SpMV / preconditioner applications: relative 1e-13 / 1e-12
  (the device sums row entries in a lane tree, the oracle sequentially);
* solves: explicit FP64 relative residual ||b - A x|| / ||b|| <= 1e-10
  (recomputed with the oracle's SpMV), and the solution error against an
  oracle solve at 1e-13 is bounded as stated in each test.
"""
import numpy as np
import pytest

import cfdrs_amd as ca
from cfdrs_amd import problems as P
from conftest import grid_laplacian, dense_to_csr
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel_res(orc, x, b):
    return np.linalg.norm(b - orc.spmv(x)) / np.linalg.norm(b)


def random_csr(n, m, row_len_max, seed, integer=False, empty_every=0):
    rng = np.random.default_rng(seed)
    rp, ci, v = [0], [], []
    for i in range(n):
        k = 0 if (empty_every and i % empty_every == 0) else int(rng.integers(0, row_len_max + 1))
        cols = np.sort(rng.choice(m, size=min(k, m), replace=False))
        vals = rng.integers(-9, 10, size=cols.size).astype(float) if integer else rng.normal(size=cols.size)
        ci.extend(cols.tolist())
        v.extend(vals.tolist())
        rp.append(len(ci))
    return np.array(rp, np.int64), np.array(ci, np.int64), np.array(v)


# m = 2777 keeps every row block's columns within 65536 (16-bit column
# offsets from a per-block base, read by the quad layout); m = 90001 does not
# (32-bit columns); "auto" takes the layout the upload rule picks
@pytest.mark.parametrize("layout", ["auto", "quad"])
@pytest.mark.parametrize("m", [2777, 90001])
@pytest.mark.parametrize("row_len", [1, 3, 7, 16, 40, 130])
def test_spmv_integer_bit_exact(row_len, m, layout):
    n = 3001
    rp, ci, v = random_csr(n, m, row_len, seed=row_len, integer=True, empty_every=97)
    x = np.random.default_rng(1).integers(-50, 51, size=m).astype(float)
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, m)
    a.set_gather(layout)
    y = a.spmv(x)
    assert np.array_equal(y, O.Csr.from_arrays(rp, ci, v, n, m).spmv(x))


# Row blocks hold at most 2048 nonzeros and 256 rows, so with 5 entries per
# row the blocks are rows [0, 256), [256, 512), ...  A block also closes
# before its columns would span more than 65535, unless some row alone does,
# and a partition stores 16-bit column offsets when every block's columns
# satisfy max - min <= 65535.  The quad and row-walk layouts read them (the
# column-sorted one keeps packed 32-bit columns), so both are forced here.
#   span65535: every row holds columns {0, 7, 1000, 30000, 65535}: 16-bit,
#     with the largest offset that fits;
#   span65536: the last column is 65536: no cut helps, 32-bit;
#   fp32_straddles: rows below 409 hold columns 0..4 and rows from 409 on
#     70000..70004, so the block from row 256 closes at row 409: 16-bit.
@pytest.mark.parametrize("layout", ["quad", "rows"])
@pytest.mark.parametrize("case", ["span65535", "span65536", "fp32_straddles"])
def test_spmv_16bit_column_span_boundary(case, layout):
    import torch

    n = 818 if case == "fp32_straddles" else 1500
    if case == "fp32_straddles":
        cols = [np.arange(5) + (0 if i < 409 else 70000) for i in range(n)]
        m = 70005
    else:
        last = 65535 if case == "span65535" else 65536
        cols = [np.array([0, 7, 1000, 30000, last])] * n
        m = last + 1
    rng = np.random.default_rng(len(case))
    ci = np.concatenate(cols).astype(np.int64)
    v = rng.integers(-9, 10, size=ci.size).astype(float)
    rp = np.arange(0, ci.size + 1, 5, dtype=np.int64)
    x = rng.integers(-50, 51, size=m).astype(float)
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, m)
    assert a.storage()["format"] == "csr_rowblock"
    assert a.set_gather(layout) == {"fp64": layout, "fp32": layout}
    yo = O.Csr.from_arrays(rp, ci, v, n, m).spmv(x)
    assert np.array_equal(a.spmv(x), yo)
    y32 = a.spmv(torch.from_numpy(x.astype(np.float32)).cuda()).cpu().numpy()
    assert np.array_equal(y32.astype(np.float64), yo)


def _integer_stencil(shape):
    nx, ny, nz = shape
    n = nx * ny * nz
    k = np.round(P.uniform(3, n, 1.0, 9.0))
    rp, ci, v, _ = P.varcoef(nx, ny, nz, 1.0, k, 0, np.zeros(6))
    v = np.round(v * 4.0)  # integer-valued, same pattern
    x = np.round(P.uniform(4, n, -20.0, 20.0))
    return n, rp, ci, v, x


# Shapes and the diagonal kernel they take (>= 4096 rows; smaller ones take
# the row-block kernel).  The FP64 two-rows-per-lane kernel needs n and the
# plane offsets nx, nx*ny to be even: (64, 64, 1), (32, 16, 12), (34, 30, 5)
# and 128^3 take it, (33, 17, 9) does not.
STENCIL_SHAPES = [(37, 1, 1), (64, 64, 1), (32, 16, 12), (34, 30, 5), (33, 17, 9), (128, 128, 128)]


@pytest.mark.parametrize("shape", STENCIL_SHAPES)
def test_spmv_stencil_integer_bit_exact(shape):
    # integer values and x make every partial sum exact
    n, rp, ci, v, x = _integer_stencil(shape)
    y = ca.CsrMatrix.from_parts(v, ci, rp, n, n).spmv(x)
    assert np.array_equal(y, O.Csr.from_arrays(rp, ci, v, n).spmv(x))


@pytest.mark.parametrize("shape", STENCIL_SHAPES[1:5])
def test_spmv_stencil_fp32_integer_bit_exact(shape):
    # |entries| <= 216 and |x| <= 20: every FP32 product and partial sum is an
    # integer below 2^24, so the FP32 product is exact as well
    import torch

    n, rp, ci, v, x = _integer_stencil(shape)
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    y32 = a.spmv(torch.from_numpy(x.astype(np.float32)).cuda()).cpu().numpy()
    assert np.array_equal(y32.astype(np.float64), O.Csr.from_arrays(rp, ci, v, n).spmv(x))


def _banded_csr(n, lo, hi, band, seed, empty_every=0):
    # rows of lo..hi integer entries within +-band of the diagonal (offsets fit
    # int16, padding small enough for the sliced-ELL layout)
    rng = np.random.default_rng(seed)
    rp, ci, v = [0], [], []
    for i in range(n):
        k = 0 if (empty_every and i % empty_every == 0) else int(rng.integers(lo, hi + 1))
        c0, c1 = max(0, i - band), min(n, i + band + 1)
        cols = np.sort(rng.choice(np.arange(c0, c1), size=min(k, c1 - c0), replace=False))
        ci.extend(cols.tolist())
        v.extend(rng.integers(-9, 10, size=cols.size).astype(float).tolist())
        rp.append(len(ci))
    return np.array(rp, np.int64), np.array(ci, np.int64), np.array(v)


def _sell_slots(rp, group):
    # slots of the sliced-ELL layout: per 64-row slice, 64 lanes x the longest
    # row rounded up to the group size
    w = np.diff(rp)
    w = np.concatenate([w, np.zeros((-w.size) % 64, w.dtype)]).reshape(-1, 64).max(axis=1)
    return int(np.sum(-(-w // group) * group) * 64)


@pytest.mark.parametrize("n,lo,hi,group", [(4096, 20, 30, 4), (5003, 20, 30, 4), (4096, 5, 5, 1), (5003, 5, 5, 1)])
def test_spmv_sell_integer_bit_exact(n, lo, hi, group):
    # sliced ELL with bounds-checked loads over the padding: uneven rows (and,
    # at 5003, empty rows and a short last slice), bit-exact in FP64 and FP32
    # on integer data; a non-finite x at a column a row does not hold must
    # not reach that row through the padding.  Rows of 20..30 entries take
    # groups of 4 slots per lane; rows of exactly 5 would be padded to 8 by
    # groups of 4, so they take groups of 1.
    import torch

    rp, ci, v = _banded_csr(n, lo, hi, 3000, seed=n, empty_every=97 if n % 2 else 0)
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    st = a.storage()
    assert st["format"] == "sell64"
    assert st["stored_entries"] == _sell_slots(rp, group)
    orc = O.Csr.from_arrays(rp, ci, v, n)
    x = np.round(P.uniform(9, n, -20.0, 20.0))
    yo = orc.spmv(x)
    assert np.array_equal(a.spmv(x), yo)
    y32 = a.spmv(torch.from_numpy(x.astype(np.float32)).cuda()).cpu().numpy()
    assert np.array_equal(y32.astype(np.float64), yo)
    x[::7] = np.inf
    y, yo = a.spmv(x), orc.spmv(x)
    assert np.array_equal(np.isfinite(y), np.isfinite(yo))
    assert np.array_equal(y[np.isfinite(yo)], yo[np.isfinite(yo)])


def test_dia_storage_selected():
    # stencil operators take the diagonal layout: the symmetric 7-point
    # operator stores its main and 3 upper diagonals, the nonsymmetric
    # 5-point upwind momentum operator all 5; scattered matrices do not
    rp, ci, v, _ = P.config2(24)
    n = 24 ** 3
    st = ca.CsrMatrix.from_parts(v, ci, rp, n, n).storage()
    assert st == {"format": "dia", "stored_entries": 4 * n, "units": 4}
    orc = _momentum()
    mrp, mci, mv = orc.arrays()
    m = mrp.size - 1
    assert ca.CsrMatrix.from_parts(mv, mci, mrp, m, m).storage()["format"] == "dia"
    assert ca.CsrMatrix.from_parts(mv, mci, mrp, m, m).storage()["units"] == 5
    rp, ci, v = random_csr(5000, 5000, 7, seed=2)
    assert ca.CsrMatrix.from_parts(v, ci, rp, 5000, 5000).storage()["format"] != "dia"


def test_spmv_dia_nonsymmetric_integer_bit_exact():
    orc = _momentum()
    rp, ci, v = orc.arrays()
    n = rp.size - 1
    # integer-valued, same (nonsymmetric) pattern, no stored zero (a stored
    # zero keeps a matrix out of diagonal storage)
    w = np.round(v * 64.0)
    v = np.where(w == 0.0, np.sign(v), w)
    assert np.all(v != 0.0)
    x = np.round(P.uniform(5, n, -30.0, 30.0))
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    assert a.storage()["format"] == "dia"
    assert np.array_equal(a.spmv(x), O.Csr.from_arrays(rp, ci, v, n).spmv(x))


def test_spmv_dia_partial_diagonals_and_absent_entries():
    # 5-point grid Laplacian on 64 x 64 with 4 % of the neighbour couplings
    # removed symmetrically: the diagonals stay >= 90 % full, so the layout is
    # kept, and the zeros it stores for absent entries must not reach y --
    # with x = inf at the end of a grid row, the next row start (one
    # diagonal step away, not a neighbour) stays finite, as in CSR
    nn = 64
    d = np.zeros((nn * nn, nn * nn))
    rng = np.random.default_rng(8)
    for i in range(nn):
        for j in range(nn):
            p = i * nn + j
            d[p, p] = 40.0
            for q in ((i + 1) * nn + j if i + 1 < nn else -1, p + 1 if j + 1 < nn else -1):
                if q >= 0 and rng.random() >= 0.04:
                    w = -float(rng.integers(1, 9))
                    d[p, q] = d[q, p] = w
    from scipy import sparse as S
    sp = S.csr_matrix(d)
    rp, ci, v = sp.indptr.astype(np.int64), sp.indices.astype(np.int64), sp.data
    n = nn * nn
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    assert a.storage() == {"format": "dia", "stored_entries": 3 * n, "units": 3}
    x = np.round(P.uniform(6, n, -30.0, 30.0))
    orc = O.Csr.from_arrays(rp, ci, v, n)
    assert np.array_equal(a.spmv(x), orc.spmv(x))
    x[nn - 1] = np.inf
    y = a.spmv(x)
    yo = orc.spmv(x)
    assert np.isfinite(y[nn]) and y[nn] == yo[nn]
    assert np.array_equal(np.isfinite(y), np.isfinite(yo))


@pytest.mark.parametrize("fmt", ["csr_rowblock", "csr_int32"])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_spmv_row_blocks_at_the_cap_bit_exact(fmt, dtype):
    # Row blocks hold at most 2048 nonzeros and are read in 4-aligned quads
    # from the aligned entry at or below their first one; the <= 3 entries
    # past the quads go to three lanes.  Rows of 2045..2048 entries after
    # short rows of 1..3 entries start blocks at every alignment and fill
    # them to the cap; 2049 and 5000 take the long-row path.  Integer data:
    # every partial sum is exact, so FP64 and FP32 must match bit for bit.
    import torch

    lens = []
    for pre in (1, 2, 3, 0):
        for big in (2045, 2046, 2047, 2048):
            lens += [pre, big] if pre else [big]
    lens += [2049, 7, 5000, 3, 1]
    rng = np.random.default_rng(17)
    m = 9000
    rp, ci, v = [0], [], []
    for L in lens:
        cols = np.sort(rng.choice(m, size=L, replace=False))
        ci.extend(cols.tolist())
        v.extend(rng.integers(-3, 4, size=L).astype(float).tolist())
        rp.append(len(ci))
    rp, ci, v = np.array(rp, np.int64), np.array(ci, np.int64), np.array(v)
    n = rp.size - 1
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, m)
    a.set_storage(fmt)
    x = np.random.default_rng(2).integers(-40, 41, size=m).astype(float)
    yo = O.Csr.from_arrays(rp, ci, v, n, m).spmv(x)
    dt = torch.float64 if dtype == "f64" else torch.float32
    y = a.spmv(torch.tensor(x, dtype=dt, device="cuda")).double().cpu().numpy()
    assert np.array_equal(y, yo)


@pytest.mark.parametrize("fmt", ["auto", "csr_rowblock", "sell64"])
def test_spmv_explicit_zero_with_non_finite_x(fmt):
    # a stored zero times inf is NaN in CSR (and in the reference's CSR
    # product); a stencil matrix that stores one is kept out of diagonal
    # storage, so every device format gives the CSR result
    nn = 64
    n = nn * nn
    rp, ci, v, _ = P.varcoef(nn, nn, 1, 1.0, np.ones(n), 0b0001, np.zeros(6))
    v = v.copy()
    v[rp[nn + 5] + 1] = 0.0  # an explicit zero in row nn + 5
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    st = a.set_storage(fmt)
    assert st["format"] != "dia"
    x = np.round(P.uniform(7, n, -30.0, 30.0))
    x[ci[rp[nn + 5] + 1]] = np.inf
    orc = O.Csr.from_arrays(rp, ci, v, n)
    y, yo = a.spmv(x), orc.spmv(x)
    assert np.isnan(yo[nn + 5]) and np.isnan(y[nn + 5])
    assert np.array_equal(np.isnan(y), np.isnan(yo))
    fin = np.isfinite(yo)
    assert np.array_equal(y[fin], yo[fin])


def test_spmv_fp32_device_tensor():
    import torch

    rp, ci, v, b = P.config2(48)
    n = 48 ** 3
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    ref = O.Csr.from_arrays(rp, ci, v, n).spmv(b.astype(np.float32).astype(np.float64))
    x32 = torch.from_numpy(b.astype(np.float32)).cuda()
    y32 = a.spmv(x32).cpu().numpy().astype(np.float64)
    # FP32 values and products: relative error per row bounded by ~nnz_row * 2^-23
    scale = np.abs(v).max() * np.abs(b).max()
    assert np.max(np.abs(y32 - ref)) <= 8 * 7 * 2.0 ** -23 * scale
    x64 = torch.from_numpy(b).cuda()
    assert np.max(np.abs(a.spmv(x64).cpu().numpy() - O.Csr.from_arrays(rp, ci, v, n).spmv(b))) <= \
        1e-13 * scale


def test_spmv_known_answer(known):
    k = known["spmv_basic"]
    a = ca.CsrMatrix.from_parts(k["values"], k["col_indices"], k["row_ptr"], 3, 3)
    assert a.spmv(np.array(k["x"])).tolist() == k["y"]
    with pytest.raises(ca.CfdlaError):
        a.spmv(np.array([1.0, 2.0]))  # dimension mismatch (sparse/tests.rs:218-219)


def test_spmv_config2_full_size_real():
    rp, ci, v, b = P.config2(128)
    n = 128 ** 3
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    y = a.spmv(b)
    yo = O.Csr.from_arrays(rp, ci, v, n).spmv(b)
    assert np.max(np.abs(y - yo)) <= 1e-13 * np.max(np.abs(yo))
    # linearity and the Neumann/Dirichlet property at full size:
    # A 1 vanishes except on the Dirichlet face z = 0, where it is k / h^2
    ones = a.spmv(np.ones(n)).reshape(128, 128, 128)
    k = P.config2_k(128).reshape(128, 128, 128)
    assert np.max(np.abs(ones[1:])) <= 1e-9 * np.max(np.abs(v))
    assert np.allclose(ones[0], k[0] * 128 * 128, rtol=1e-12)
    y2 = a.spmv(2.0 * b + ones.ravel())
    assert np.allclose(y2, 2.0 * y + a.spmv(ones.ravel()), rtol=1e-12, atol=1e-9 * np.max(np.abs(y)))


def test_jacobi_apply_matches_oracle():
    rp, ci, v, b = P.config1(64)
    n = 64 * 64
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    z = np.empty(n)
    ca.JacobiPreconditioner(a).apply_to(b, z)
    assert np.array_equal(z, b * O.jacobi_inv_diag(O.Csr.from_arrays(rp, ci, v, n)))


AMG_CASES = [
    dict(coarsening="aggregation", interp="direct", coarse="direct", omega=0.8),
    dict(coarsening="aggregation", interp="direct", coarse="sweeps", omega=1.0),
    dict(coarsening="three_phase", interp="smoothed", coarse="direct", omega=0.8),
    dict(coarsening="three_phase", interp="direct", coarse="sweeps", omega=0.7),
]


# nn = 20 runs the fine level on the two-rows-per-lane diagonal kernel,
# nn = 19 on the one-row-per-lane kernel
@pytest.mark.parametrize("nn", [20, 19])
@pytest.mark.parametrize("case", AMG_CASES, ids=lambda c: f"{c['coarsening']}-{c['interp']}-{c['coarse']}")
def test_amg_vcycle_matches_oracle(case, nn):
    n = nn ** 3
    k = P.config2_k(nn)
    rp, ci, v, _ = P.varcoef(nn, nn, nn, float(nn * nn), k, 1 << 4, np.zeros(6))
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    smoothed = case["interp"] == "smoothed"
    if case["coarsening"] == "aggregation":
        cfg = ca.AMGConfig.gpu_aggregation(smoothed=smoothed, relaxation_factor=case["omega"])
    else:
        cfg = ca.AMGConfig.smoothed_aggregation(0.25, case["omega"], smoothed=smoothed)
    cfg.coarse_solver = ca.CoarseSolver.Direct if case["coarse"] == "direct" else ca.CoarseSolver.ReferenceSweeps
    amg = ca.AlgebraicMultigrid(a, cfg)
    oamg = O.Amg(O.Csr.from_arrays(rp, ci, v, n), coarsening=case["coarsening"], interp=case["interp"],
                 smoother="jacobi", omega=case["omega"], coarse_solve=case["coarse"], sa_omega=cfg.smoothing_omega,
                 theta=cfg.strength_threshold, min_coarse=cfg.min_coarse_size)
    r = P.uniform(77, n)
    z = np.empty(n)
    amg.apply_to(r, z)
    zo = oamg.apply(r)
    assert np.linalg.norm(z - zo) <= 1e-11 * np.linalg.norm(zo)


@pytest.mark.parametrize("nn", [24, 20])
def test_fused_coarse_restriction_matches_separate_products(nn):
    # the level above the dense coarsest one runs its two Jacobi pre-sweeps,
    # residual and restriction as one product with [2wD - w^2 DAD; R (I - wAD)^2]:
    # the same V-cycle as the three products up to rounding, and the oracle's
    import torch

    n = nn ** 3
    rp, ci, v, _ = P.varcoef(nn, nn, nn, float(nn * nn), P.config2_k(nn), 1 << 4, np.zeros(6))
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    cfg = ca.AMGConfig.smoothed_aggregation()
    cfg.min_coarse_size = 100
    sep = ca.AMGConfig(**{**cfg.__dict__, "separate_coarse_restriction": True})
    fused, plain = ca.AlgebraicMultigrid(a, cfg), ca.AlgebraicMultigrid(a, sep)
    r = P.uniform(78, n)
    z, zs = np.empty(n), np.empty(n)
    fused.apply_to(r, z)
    plain.apply_to(r, zs)
    L = fused.statistics.num_levels
    assert L >= 3
    assert fused.statistics.fused_level == L - 2 and plain.statistics.fused_level == -1
    assert np.linalg.norm(z - zs) <= 1e-13 * np.linalg.norm(zs)
    oamg = O.Amg(O.Csr.from_arrays(rp, ci, v, n), coarsening="three_phase", interp="smoothed", smoother="jacobi",
                 omega=cfg.relaxation_factor, coarse_solve="direct", sa_omega=cfg.smoothing_omega,
                 theta=cfg.strength_threshold, min_coarse=cfg.min_coarse_size)
    zo = oamg.apply(r)
    assert np.linalg.norm(z - zo) <= 1e-11 * np.linalg.norm(zo)
    rt = torch.tensor(r, dtype=torch.float32, device="cuda")
    z32, zs32 = torch.empty_like(rt), torch.empty_like(rt)
    fused.apply_to(rt, z32)
    plain.apply_to(rt, zs32)
    z32, zs32 = z32.double().cpu().numpy(), zs32.double().cpu().numpy()
    assert np.linalg.norm(z32 - zo) <= 1e-5 * np.linalg.norm(zo)
    assert np.linalg.norm(z32 - zs32) <= 1e-5 * np.linalg.norm(zs32)


def test_pcg_jacobi_config1():
    rp, ci, v, b = P.config1(128)
    n = 128 * 128
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    orc = O.Csr.from_arrays(rp, ci, v, n)
    x = np.zeros(n)
    cfg = ca.IterativeSolverConfig.new(1e-10).with_max_iterations(20000)
    rep = ca.cg_preconditioned(a, b, ca.JacobiPreconditioner(a), x, cfg)
    assert rep.termination == ca.Termination.Converged
    assert rel_res(orc, x, b) <= 1e-10
    xo, orep = O.solve(orc, b, kind="cg", prec="jacobi", tol_rel=1e-13, max_it=50000)
    err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    assert err <= 1e-5, err  # kappa(A) ~ 1e4 times the 1e-10 residual
    # iteration counts agree with the oracle's textbook PCG to within a few steps
    _, orep10 = O.solve(orc, b, kind="cg", prec="jacobi", tol_rel=1e-10, max_it=50000)
    assert abs(rep.iterations - orep10.iterations) <= max(3, 0.02 * orep10.iterations)


def _config2_small(nn=32):
    n = nn ** 3
    k = P.config2_k(nn)
    rp, ci, v, _ = P.varcoef(nn, nn, nn, float(nn * nn), k, 1 << 4, np.zeros(6))
    b = P.uniform(P.SEED_C2_B, n)
    return rp, ci, v, b, n


@pytest.mark.parametrize("precision", [ca.Precision.FP64, ca.Precision.FP32_REFINED])
@pytest.mark.parametrize("recipe", ["greedy-direct", "three-phase-smoothed"])
def test_amg_pcg_config2_small(precision, recipe):
    rp, ci, v, b, n = _config2_small(32)
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    orc = O.Csr.from_arrays(rp, ci, v, n)
    cfg = ca.AMGConfig.gpu_aggregation(smoothed=False) if recipe == "greedy-direct" else \
        ca.AMGConfig.smoothed_aggregation()
    amg = ca.AlgebraicMultigrid(a, cfg)
    x = np.zeros(n)
    cfg = ca.IterativeSolverConfig.new(1e-10).with_max_iterations(2000).with_precision(precision)
    rep = ca.cg_preconditioned(a, b, amg, x, cfg)
    assert rep.termination == ca.Termination.Converged, rep
    assert rel_res(orc, x, b) <= 1e-10
    xo, _ = O.solve(orc, b, kind="cg", prec="jacobi", tol_rel=1e-13, max_it=100000)
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-6


def _momentum(nn=64, seed=3):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1, 1, nn * nn)
    w = rng.uniform(-1, 1, nn * nn)
    a = O.assemble_momentum_u(nn, nn, 1 / (nn - 1), 1 / (nn - 1), 1e-2, 1e-3, 1.0, u, w)
    return a


@pytest.mark.parametrize("kind", ["bicgstab", "gmres"])
@pytest.mark.parametrize("precision,reliable", [(ca.Precision.FP64, 0), (ca.Precision.FP32_REFINED, 0),
                                                (ca.Precision.FP32_REFINED, -1)],
                         ids=["fp64", "fp32-reliable", "fp32-restarted"])
def test_nonsymmetric_solvers(kind, precision, reliable):
    # FP32 BiCGSTAB runs with FP64 residual replacement (reliable 0) or
    # restarted refinement (-1); GMRES always restarts
    orc = _momentum()
    rp, ci, v = orc.arrays()
    n = rp.size - 1
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    b = P.uniform(31, n)
    x = np.zeros(n)
    cfg = (ca.IterativeSolverConfig.new(1e-10).with_max_iterations(5000).with_precision(precision)
           .with_reliable_interval(reliable))
    jac = ca.JacobiPreconditioner(a)
    if kind == "bicgstab":
        rep = ca.bicgstab_preconditioned(a, b, jac, x, cfg)
    else:
        rep = ca.gmres_preconditioned(a, b, jac, x, cfg, 30)
    assert rep.termination == ca.Termination.Converged, rep
    assert rel_res(orc, x, b) <= 1e-10
    xo = np.linalg.solve(orc.dense(), b)
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-8


def test_gmres_diagonal_known_answer(known):
    k = known["gmres_diagonal"]
    a = ca.CsrMatrix.from_parts(k["diag"], [0, 1], [0, 1, 2], 2, 2)
    ws = ca.KrylovWorkspace(k["restart"], 2)
    cfg = ca.IterativeSolverConfig.new(k["tolerance"]).with_max_iterations(k["max_iterations"])
    for _ in range(2):
        x = np.zeros(2)
        rep = ca.gmres_preconditioned_with_workspace(a, np.array(k["b"]), None, x, cfg, ws)
        assert rep.final_residual_norm <= rep.threshold
        assert np.allclose(x, k["x"], rtol=0, atol=k["solution_tol"])


@pytest.mark.parametrize("kind", ["cg", "bicgstab", "gmres"])
def test_gmres_matches_oracle_iterations_exact_arithmetic_cases(kind):
    # tiny SPD system: every method terminates within n steps (+ restarts)
    rng = np.random.default_rng(12)
    m = rng.normal(size=(10, 10))
    d = m @ m.T + 10 * np.eye(10)
    rp, ci, v = dense_to_csr(d)
    a = ca.CsrMatrix.from_parts(v, ci, rp, 10, 10)
    b = rng.normal(size=10)
    x = np.zeros(10)
    cfg = ca.IterativeSolverConfig.new(1e-12).with_max_iterations(40)
    rep = ca.SolverKind.Gmres(16).solve(a, b, x, cfg) if kind == "gmres" else (
        ca.cg(a, b, x, cfg) if kind == "cg" else ca.bicgstab(a, b, x, cfg))
    assert rep.termination == ca.Termination.Converged
    assert np.allclose(x, np.linalg.solve(d, b), rtol=1e-9, atol=1e-11)


def test_initial_residual_and_stall_and_curvature():
    rp, ci, v, b = P.config1(32)
    n = 1024
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    x = np.zeros(n)
    rep = ca.cg(a, np.zeros(n), x, ca.IterativeSolverConfig.new(1e-10))
    assert rep.termination == ca.Termination.InitialResidual and rep.iterations == 0
    x = np.zeros(n)
    rep = ca.cg(a, b, x, ca.IterativeSolverConfig.new(1e-12).with_max_iterations(3))
    assert rep.termination == ca.Termination.MaxIterations and rep.iterations == 3
    assert ca.interpret("p", rep).kind == ca.SolveOutcome.Stalled
    assert np.any(x != 0)  # the last iterate is kept
    d = ca.CsrMatrix.from_parts([1.0, -1.0, 2.0], [0, 1, 2], [0, 1, 2, 3], 3, 3)
    x = np.zeros(3)
    rep = ca.cg(d, np.ones(3), x, ca.IterativeSolverConfig.new(1e-12).with_max_iterations(10))
    assert rep.termination in (ca.Termination.NonPositiveCurvature, ca.Termination.Breakdown)
    assert ca.interpret("p", rep).kind == ca.SolveOutcome.BrokenDown


def test_dimension_errors():
    rp, ci, v, b = P.config1(8)
    a = ca.CsrMatrix.from_parts(v, ci, rp, 64, 64)
    with pytest.raises(ca.SolveError):
        ca.cg(a, np.ones(63), np.zeros(63), ca.IterativeSolverConfig.new(1e-8))
    small = ca.CsrMatrix.from_parts([1.0], [0], [0, 1], 1, 1)
    with pytest.raises(ca.SolveError):
        ca.cg_preconditioned(a, b, ca.JacobiPreconditioner(small), np.zeros(64), ca.IterativeSolverConfig.new(1e-8))


def test_solves_are_bitwise_reproducible():
    rp, ci, v, b, n = _config2_small(24)
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    amg = ca.AlgebraicMultigrid(a, ca.AMGConfig.smoothed_aggregation())
    xs = []
    for _ in range(2):
        x = np.zeros(n)
        ca.cg_preconditioned(a, b, amg, x, ca.IterativeSolverConfig.new(1e-10).with_max_iterations(500))
        xs.append(x)
    assert np.array_equal(xs[0], xs[1])


def test_amg_recompute_then_solve():
    rp, ci, v, b, n = _config2_small(16)
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    amg = ca.AlgebraicMultigrid(a, ca.AMGConfig.smoothed_aggregation())
    v2 = v * 3.0
    a.update_values(v2)
    amg.recompute(ca.CsrMatrix.from_parts(v2, ci, rp, n, n))
    x = np.zeros(n)
    rep = ca.cg_preconditioned(a, b, amg, x, ca.IterativeSolverConfig.new(1e-10).with_max_iterations(500))
    assert rep.termination == ca.Termination.Converged
    assert rel_res(O.Csr.from_arrays(rp, ci, v2, n), x, b) <= 1e-10


@pytest.mark.parametrize("interval,delta", [(0, 0.0), (0, 1e-1), (0, 1e-5), (1, 0.0), (3, 0.0), (8, 0.0),
                                            (32, 0.0), (-1, 0.0)])
def test_refined_cg_reliable_updates_amg(interval, delta):
    # FP32 PCG with FP64 residual replacement when the residual dropped by
    # delta (interval 0) or every `interval` iterations (-1: restarted
    # refinement).  All reach the FP64 threshold; the uninterrupted
    # recurrence keeps the FP64 iteration count.  interval 32 replaces only at
    # convergence, where the FP64 residual is far above the recursive one:
    # the direction restart keeps it from diverging.
    rp, ci, v, b, n = _config2_small(32)
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    orc = O.Csr.from_arrays(rp, ci, v, n)
    amg = ca.AlgebraicMultigrid(a, ca.AMGConfig.smoothed_aggregation())
    x64 = np.zeros(n)
    r64 = ca.cg_preconditioned(a, b, amg, x64, ca.IterativeSolverConfig.new(1e-10).with_max_iterations(500))
    x = np.zeros(n)
    cfg = (ca.IterativeSolverConfig.new(1e-10).with_max_iterations(500)
           .with_precision(ca.Precision.FP32_REFINED).with_reliable_interval(interval).with_reliable_delta(delta))
    rep = ca.cg_preconditioned(a, b, amg, x, cfg)
    assert rep.termination == ca.Termination.Converged, rep
    res = rel_res(orc, x, b)
    assert res <= 1e-10
    # the reported norm is the explicit FP64 residual of the returned x
    assert rep.final_residual_norm == pytest.approx(res * np.linalg.norm(b), rel=1e-6)
    assert np.linalg.norm(x - x64) / np.linalg.norm(x64) <= 1e-8
    if 0 <= interval <= 8 and delta != 1e-5:
        assert rep.iterations <= r64.iterations + 2, (rep.iterations, r64.iterations)
    if interval >= 0:
        assert rep.iterations <= 2 * r64.iterations, (rep.iterations, r64.iterations)
        assert rep.outer_iterations >= 1


def test_refined_cg_reliable_updates_jacobi_config1():
    # kappa ~ 1e4: the FP32 recurrence loses orthogonality faster than the
    # FP64 one (measured 909 against 750 iterations; restarted refinement
    # takes 1471)
    rp, ci, v, b = P.config1(128)
    n = 128 * 128
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    orc = O.Csr.from_arrays(rp, ci, v, n)
    jac = ca.JacobiPreconditioner(a)
    x64 = np.zeros(n)
    r64 = ca.cg_preconditioned(a, b, jac, x64, ca.IterativeSolverConfig.new(1e-10).with_max_iterations(20000))
    x = np.zeros(n)
    cfg = ca.IterativeSolverConfig.new(1e-10).with_max_iterations(20000).with_precision(ca.Precision.FP32_REFINED)
    rep = ca.cg_preconditioned(a, b, jac, x, cfg)
    assert rep.termination == ca.Termination.Converged, rep
    assert rel_res(orc, x, b) <= 1e-10
    assert rep.iterations <= 1.3 * r64.iterations, (rep.iterations, r64.iterations)


def test_refined_bicgstab_reliable_keeps_fp64_iteration_count():
    orc = _momentum(nn=128, seed=5)
    rp, ci, v = orc.arrays()
    n = rp.size - 1
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    jac = ca.JacobiPreconditioner(a)
    b = P.uniform(33, n)
    reps = {}
    for name, prec, rel in (("fp64", ca.Precision.FP64, 0), ("reliable", ca.Precision.FP32_REFINED, 0),
                            ("restarted", ca.Precision.FP32_REFINED, -1)):
        x = np.zeros(n)
        cfg = (ca.IterativeSolverConfig.new(1e-10).with_max_iterations(20000).with_precision(prec)
               .with_reliable_interval(rel))
        reps[name] = ca.bicgstab_preconditioned(a, b, jac, x, cfg)
        assert reps[name].termination == ca.Termination.Converged, (name, reps[name])
        assert rel_res(orc, x, b) <= 1e-10, name
    assert reps["reliable"].iterations <= 1.3 * reps["fp64"].iterations + 5, \
        (reps["reliable"].iterations, reps["fp64"].iterations, reps["restarted"].iterations)


def test_refined_cg_reliable_budget_and_initial_residual():
    rp, ci, v, b = P.config1(32)
    n = 1024
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    orc = O.Csr.from_arrays(rp, ci, v, n)
    cfg = ca.IterativeSolverConfig.new(1e-12).with_max_iterations(5).with_precision(ca.Precision.FP32_REFINED)
    x = np.zeros(n)
    rep = ca.cg_preconditioned(a, b, ca.JacobiPreconditioner(a), x, cfg)
    assert rep.termination == ca.Termination.MaxIterations and rep.iterations == 5
    # the FP32 correction is folded into x and the reported norm is its FP64 residual
    assert np.any(x != 0)
    assert rep.final_residual_norm == pytest.approx(np.linalg.norm(b - orc.spmv(x)), rel=1e-9)
    x = np.zeros(n)
    rep = ca.cg_preconditioned(a, np.zeros(n), ca.JacobiPreconditioner(a), x, cfg)
    assert rep.termination == ca.Termination.InitialResidual and rep.iterations == 0
    d = ca.CsrMatrix.from_parts([1.0, -1.0, 2.0], [0, 1, 2], [0, 1, 2, 3], 3, 3)
    x = np.zeros(3)
    rep = ca.cg(d, np.ones(3), x, ca.IterativeSolverConfig.new(1e-12).with_max_iterations(10)
                .with_precision(ca.Precision.FP32_REFINED))
    assert rep.termination in (ca.Termination.NonPositiveCurvature, ca.Termination.Breakdown)


def test_sor_diagonal_closed_form():
    # one colour: z = omega r / d
    d = np.array([2.0, 4.0, 8.0, 5.0])
    a = ca.CsrMatrix.from_parts(d, [0, 1, 2, 3], [0, 1, 2, 3, 4], 4, 4)
    sor = ca.SuccessiveOverRelaxation(a, 1.5)
    assert sor.colours == 1
    r = np.array([1.0, -2.0, 3.0, 0.5])
    z = np.empty(4)
    sor.apply_to(r, z)
    assert np.array_equal(z, (1.5 / d) * r)


def test_sor_red_black_tridiagonal_by_hand():
    # 1D operator: rows coloured even / odd; the even sweep sees z = 0, the
    # odd sweep the even z -- written out with the two recurrences
    n, w = 9, 1.2
    main, lo, up = np.full(n, 4.0), -np.arange(1, n) * 0.5, -np.ones(n - 1)
    dense = np.diag(main) + np.diag(lo, -1) + np.diag(up, 1)
    rp, ci, v = dense_to_csr(dense)
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    sor = ca.SuccessiveOverRelaxation(a, w)
    assert sor.colours == 2
    r = np.linspace(-1.0, 2.0, n)
    z = np.empty(n)
    sor.apply_to(r, z)
    expect = np.zeros(n)
    for i in range(0, n, 2):
        expect[i] = (w / main[i]) * r[i]
    for i in range(1, n, 2):
        acc = r[i] - lo[i - 1] * expect[i - 1]
        if i + 1 < n:
            acc -= up[i] * expect[i + 1]
        expect[i] = (w / main[i]) * acc
    assert np.allclose(z, expect, rtol=1e-15, atol=0)


def test_sor_dense_is_lexicographic_sweep():
    # every pair of rows is coupled, so each row takes its own colour in
    # visit order 0, 1, ..., n-1: the sweep is the sequential forward SOR
    # sweep from zero, z = (D / omega + L)^-1 r with L the strict lower part
    n, w = 7, 1.3
    rng = np.random.default_rng(11)
    dense = rng.uniform(-1.0, 1.0, (n, n)) + np.diag(np.full(n, 6.0))
    rp, ci, v = dense_to_csr(dense)
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    sor = ca.SuccessiveOverRelaxation(a, w)
    assert sor.colours == n
    r = rng.uniform(-1.0, 1.0, n)
    z = np.empty(n)
    sor.apply_to(r, z)
    expect = np.linalg.solve(np.diag(np.diag(dense)) / w + np.tril(dense, -1), r)
    assert np.allclose(z, expect, rtol=1e-13, atol=1e-15)


def test_sor_unit_pivot_without_diagonal():
    # row 1 carries no self-coupling: unit pivot, as from_csr_with_identity_rows
    a = ca.CsrMatrix.from_parts([2.0, 1.0, 3.0], [0, 0, 2], [0, 1, 2, 3], 3, 3)
    sor = ca.SuccessiveOverRelaxation.from_csr_with_identity_rows(a, 1.0)
    r = np.array([4.0, 1.0, 6.0])
    z = np.empty(3)
    sor.apply_to(r, z)
    # colours: row 0 -> 0, row 1 (coupled to 0) -> 1, row 2 -> 0
    assert np.allclose(z, [2.0, 1.0 - 1.0 * 2.0, 2.0], rtol=1e-15, atol=0)


@pytest.mark.parametrize("kind", ["gmres", "bicgstab"])
@pytest.mark.parametrize("precision", [ca.Precision.FP64, ca.Precision.FP32_REFINED])
def test_sor_preconditioned_momentum(kind, precision):
    # the reference's momentum preconditioner (SOR, omega = 1) in red-black
    # order: converges to the dense solution in fewer steps than Jacobi
    orc = _momentum()
    rp, ci, v = orc.arrays()
    n = rp.size - 1
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    b = P.uniform(31, n)
    sor = ca.SuccessiveOverRelaxation(a, 1.0)
    assert sor.colours == 2
    cfg = ca.IterativeSolverConfig.new(1e-10).with_max_iterations(5000).with_precision(precision)
    its = {}
    for name, m in (("sor", sor), ("jacobi", ca.JacobiPreconditioner(a))):
        x = np.zeros(n)
        rep = (ca.gmres_preconditioned(a, b, m, x, cfg, 30) if kind == "gmres"
               else ca.bicgstab_preconditioned(a, b, m, x, cfg))
        assert rep.termination == ca.Termination.Converged, (name, rep)
        assert rel_res(orc, x, b) <= 1e-10
        its[name] = rep.iterations
        if name == "sor":
            xo = np.linalg.solve(orc.dense(), b)
            assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-8
    assert its["sor"] < its["jacobi"], its
