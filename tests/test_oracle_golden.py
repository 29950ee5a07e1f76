"""Pins the CPU oracle against the reference's own known answers and against
what the mathematics fixes (dense products, exact termination, closed forms).
CPU only."""
import math

import numpy as np
import pytest

from conftest import dense_to_csr, grid_laplacian, tridiag
from oracle import oracle as O


def csr(d):
    return O.Csr.from_dense(np.asarray(d, float))


# ---- smoothers (smoothers.rs:383-491) --------------------------------------------
@pytest.mark.parametrize("kind,key", [(O.SM_GS, "gauss_seidel"), (O.SM_JACOBI, "jacobi"),
                                      (O.SM_SGS, "symmetric_gauss_seidel"), (O.SM_SOR, "sor")])
def test_smoother_known_answers(known, kind, key):
    k = known["smoothers"]
    x = O.smoother(csr(tridiag(5)), kind, k["omega"], k["x0"], k["b"], k["sweeps"])
    assert x.tolist() == k[key]


def test_chebyshev_bounds_and_first_entry(known):
    a = csr(tridiag(5))
    lo, hi = O.chebyshev_bounds(a)
    assert [lo, hi] == known["smoothers"]["chebyshev_bounds"]
    b = np.array(known["smoothers"]["b"])
    x = O.smoother(a, O.SM_CHEBYSHEV, 1.0, np.zeros(5), b, 1)
    theta, delta = (hi + lo) / 2, (hi - lo) / 2
    sigma = theta / delta
    rho_old = 1 / sigma
    alpha = (1 / (2 * sigma - rho_old)) / rho_old
    assert abs(x[0] - alpha * b[0] / 2.0) <= np.finfo(float).eps * 16


# ---- strength / coarsening / interpolation -------------------------------------------
def test_strength_matrix_4x4(known):
    k = known["strength_4x4"]
    s = O.strength(csr(grid_laplacian(4)), k["theta"])
    rp, ci, v = s.arrays()
    assert s.shape[2] == k["nnz"]
    assert ci[rp[0]:rp[1]].tolist() == k["row0_cols"] and v[rp[0]:rp[1]].tolist() == k["row0_vals"]
    assert ci[rp[5]:rp[6]].tolist() == k["row5_cols"] and v[rp[5]:rp[6]].tolist() == k["row5_vals"]


def test_strength_brute_force_random():
    rng = np.random.default_rng(7)
    for _ in range(20):
        d = rng.uniform(-1, 1, (12, 12)) * (rng.random((12, 12)) < 0.4)
        np.fill_diagonal(d, 5.0)
        s = O.strength(csr(d), 0.5).dense()
        for i in range(12):
            off = [abs(d[i, j]) for j in range(12) if j != i and d[i, j] != 0]
            mx = max(off) if off else 0.0
            for j in range(12):
                strong = j != i and d[i, j] != 0 and abs(d[i, j]) >= 0.5 * mx
                assert (s[i, j] != 0) == strong


@pytest.mark.parametrize("kind", ["rs", "aggregation", "hybrid"])
def test_coarsening_mapping_invariants(kind):
    # coarsening/mod.rs:134-181: mapped indices in range; C points map to themselves
    a = csr(grid_laplacian(4))
    cp, f2c = O.coarsen(a, kind, 0.25, 4)
    for i, m in enumerate(f2c):
        if m >= 0:
            assert m < len(cp)
            if i in cp.tolist():
                assert cp[m] == i
    for idx, c in enumerate(cp):
        assert f2c[c] == idx


def test_aggregation_greedy_properties():
    # every point aggregated; each aggregate = root + strong neighbours of the root; size <= max
    a = csr(grid_laplacian(10))
    cp, f2c = O.coarsen(a, "aggregation", 0.5, 8)
    assert (f2c >= 0).all()
    s = O.strength(a, 0.5).dense()
    for agg, root in enumerate(cp):
        members = np.nonzero(f2c == agg)[0]
        assert root == members.min() and len(members) <= 8
        for m in members:
            assert m == root or s[root, m] != 0


def test_rs_coarsening_ratio():
    # coarsening/mod.rs:185-198 (ratio bounds on a 10x10 grid)
    a = csr(grid_laplacian(10))
    cp, f2c = O.coarsen(a, "rs", 0.25)
    assert 0.15 <= len(cp) / 100 <= 0.65
    assert (f2c >= 0).all()


def test_direct_interpolation(known):
    k = known["direct_interpolation"]
    p = O.direct_interp(np.array(k["fine_to_coarse"]), 5, 3).dense()
    for r, c, v in k["entries"]:
        assert p[r, c] == v
    for r, c in k["zeros"]:
        assert p[r, c] == 0.0


def test_classical_interpolation_1d(known):
    k = known["classical_interpolation_1d"]
    s = np.abs(tridiag(5, 0.0, 1.0))
    p = O.classical_interp(csr(tridiag(5)), k["coarse_points"], csr(s)).dense()
    assert abs(p[1, 0] - k["w10"]) < 1e-10 and abs(p[1, 1] - k["w11"]) < 1e-10
    assert p[0, 0] == 1.0 and p[2, 1] == 1.0 and p[4, 2] == 1.0


def test_classical_interpolation_ff_weights(known):
    k = known["classical_interpolation_ff"]
    n = k["n"]
    d = np.eye(n) * k["diag"]
    for u, v in k["edges"]:
        d[u, v] = d[v, u] = -1.0
    s = (np.abs(d) > 0).astype(float)
    np.fill_diagonal(s, 0.0)
    p = O.classical_interp(csr(d), k["coarse_points"], csr(s)).dense()
    assert abs(p[1, 0] - k["w1_0"]) < 1e-10 and abs(p[1, 1] - k["w1_3"]) < 1e-10


def test_transpose_and_galerkin_match_dense():
    rng = np.random.default_rng(3)
    for _ in range(10):
        a = rng.uniform(-1, 1, (15, 15)) * (rng.random((15, 15)) < 0.3) + np.eye(15) * 4
        p = rng.uniform(0, 1, (15, 6)) * (rng.random((15, 6)) < 0.4)
        A, P = csr(a), csr(p)
        R = P.transpose()
        assert np.array_equal(R.dense(), p.T)
        Ac = (R @ A) @ P
        assert np.allclose(Ac.dense(), p.T @ a @ p, rtol=1e-13, atol=1e-13)


def test_smoothed_prolongator_matches_dense():
    a = grid_laplacian(6)
    A = csr(a)
    cp, f2c = O.coarsen(A, "aggregation", 0.5, 8)
    pt = O.direct_interp(f2c, 36, len(cp))
    p = O.smooth_interp(A, pt, 2.0 / 3.0).dense()
    expect = (np.eye(36) - (2.0 / 3.0) * np.diag(1 / np.diag(a)) @ a) @ pt.dense()
    assert np.allclose(p, expect, rtol=0, atol=1e-14)


# ---- SpMV and Krylov --------------------------------------------------------------------
def test_spmv_basic(known):
    k = known["spmv_basic"]
    a = O.Csr.from_arrays(k["row_ptr"], k["col_indices"], k["values"], 3)
    assert a.spmv(np.array(k["x"])).tolist() == k["y"]


def test_restart_ladder(known):
    for req, width in known["restart_ladder"]["cases"]:
        assert O.restart_ladder(req) == width


@pytest.mark.parametrize("kind", ["cg", "bicgstab", "gmres"])
def test_diagonal_system(known, kind):
    k = known["gmres_diagonal"]
    a = csr(np.diag(k["diag"]))
    x, rep = O.solve(a, k["b"], kind=kind, tol_rel=k["tolerance"], max_it=k["max_iterations"],
                     restart=k["restart"])
    assert rep.final_residual_norm <= rep.threshold
    assert np.allclose(x, k["x"], rtol=0, atol=k["solution_tol"])


def test_cg_exact_termination_small_spd():
    # CG terminates in <= n steps in exact arithmetic; FP64 meets 1e-12 within n+2
    rng = np.random.default_rng(11)
    m = rng.normal(size=(8, 8))
    a = m @ m.T + 8 * np.eye(8)
    b = rng.normal(size=8)
    x, rep = O.solve(csr(a), b, kind="cg", tol_rel=1e-12, max_it=10)
    assert rep.termination == O.T_CONVERGED and rep.iterations <= 10
    assert np.allclose(x, np.linalg.solve(a, b), rtol=1e-10, atol=1e-12)


def test_gmres_full_restart_is_exact():
    rng = np.random.default_rng(5)
    a = rng.normal(size=(12, 12)) + 6 * np.eye(12)
    b = rng.normal(size=12)
    x, rep = O.solve(csr(a), b, kind="gmres", tol_rel=1e-13, max_it=12, restart=16)
    assert rep.termination == O.T_CONVERGED
    assert np.allclose(x, np.linalg.solve(a, b), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("kind", ["cg", "bicgstab", "gmres"])
@pytest.mark.parametrize("prec", ["identity", "jacobi"])
def test_poisson_solutions_match_dense(kind, prec):
    # cfd-validation/src/numerical/linear_solver.rs:52-220 style: relative solution error
    a = grid_laplacian(8)
    b = np.sin(np.arange(64.0))
    x, rep = O.solve(csr(a), b, kind=kind, prec=prec, tol_rel=1e-12, max_it=500, restart=30)
    assert rep.termination == O.T_CONVERGED
    xs = np.linalg.solve(a, b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-10


def test_cg_nonpositive_curvature_on_indefinite():
    a = np.diag([1.0, -1.0, 2.0])
    x, rep = O.solve(csr(a), np.ones(3), kind="cg", tol_rel=1e-12, max_it=10)
    assert rep.termination in (O.T_NONPOS, O.T_BREAKDOWN)


def test_stall_reports_max_iterations():
    a = grid_laplacian(8)
    x, rep = O.solve(csr(a), np.ones(64), kind="cg", tol_rel=1e-14, max_it=1)
    assert rep.termination == O.T_MAXIT and rep.iterations == 1


# ---- AMG (amg_integration_test.rs:97-268) ---------------------------------------------
def _manufactured(n):
    t = np.arange(n * n) / (n * n)
    return np.sin(np.pi * t) + 0.5 * t


@pytest.mark.parametrize("kind", ["bicgstab", "gmres"])
def test_amg_default_manufactured_solution(known, kind):
    k = known["amg_manufactured"]
    n = k["grid_n"]
    a = grid_laplacian(n)
    A = csr(a)
    xs = _manufactured(n)
    b = a @ xs
    amg = O.Amg(A)  # AMGConfig::default(): RS + classical + Gauss-Seidel
    x, rep = O.solve(A, b, kind=kind, prec="amg", amg=amg, tol_rel=k["tolerance"], max_it=k["max_iterations"],
                     restart=30)
    assert rep.termination == O.T_CONVERGED
    assert np.max(np.abs(x - xs)) < k["max_abs_error"]


@pytest.mark.parametrize("n", [8, 16])
def test_amg_stationary_contraction(n):
    # amg_integration_test.rs:211-268: x <- x + M^-1 (b - A x) contracts in the A-norm
    a = grid_laplacian(n)
    A = csr(a)
    xs = _manufactured(n)
    b = a @ xs
    for cfg in (dict(), dict(coarsening="aggregation", interp="smoothed", smoother="jacobi", omega=0.8,
                             coarse_solve="direct")):
        amg = O.Amg(A, **cfg)
        x = np.zeros(n * n)
        e0 = math.sqrt((x - xs) @ a @ (x - xs))
        for _ in range(3):
            x = x + amg.apply(b - a @ x)
        e1 = math.sqrt((x - xs) @ a @ (x - xs))
        assert (e1 / e0) ** (1 / 3) < 1.0


def test_jacobi_vcycle_is_symmetric():
    # Jacobi pre = post and R = P^T: z^T M^-1 y = y^T M^-1 z
    a = grid_laplacian(12)
    amg = O.Amg(csr(a), coarsening="aggregation", interp="smoothed", smoother="jacobi", omega=0.8,
                coarse_solve="direct")
    rng = np.random.default_rng(2)
    y, z = rng.normal(size=144), rng.normal(size=144)
    lhs, rhs = z @ amg.apply(y), y @ amg.apply(z)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


def test_galerkin_levels_match_dense_triple_product():
    a = grid_laplacian(10)
    amg = O.Amg(csr(a), coarsening="aggregation", interp="direct", smoother="jacobi", min_coarse=10)
    for l in range(amg.num_levels - 1):
        A, P = amg.level(l, "A").dense(), amg.level(l, "P").dense()
        assert np.allclose(amg.level(l + 1, "A").dense(), P.T @ A @ P, rtol=1e-13, atol=1e-12)


# ---- 1D networks (matrix_assembly.rs:335-455, linear_system.rs:339-349) --------------------
def _net(edges, dirichlet, n, neumann=None):
    ei = np.array([e[0] for e in edges])
    ej = np.array([e[1] for e in edges])
    g = np.array([e[2] for e in edges], float)
    is_dir = np.zeros(n, np.uint8)
    dv = np.zeros(n)
    for k, v in dirichlet.items():
        is_dir[int(k)] = 1
        dv[int(k)] = v
    return O.assemble_network(n, ei, ej, g, is_dir, dv, neumann)


def test_network_three_node(known):
    k = known["network_three_node"]
    a, rhs = _net(k["edges"], k["dirichlet"], 3)
    d = a.dense()
    assert d[0, 0] == 1.0 and d[2, 2] == 1.0 and d[0, 1] == 0.0 and d[2, 1] == 0.0
    assert d[1, 1] == k["diag_b"] and d[1, 0] == 0.0 and d[1, 2] == 0.0
    assert rhs.tolist() == k["rhs"]
    neu = np.zeros(3)
    neu[1] = k["neumann_b"]
    _, rhs2 = _net(k["edges"], k["dirichlet"], 3, neu)
    assert rhs2[1] == k["rhs_with_neumann_b"]


def test_network_active_block_symmetric(known):
    k = known["network_symmetric_block"]
    a, _ = _net(k["edges"], k["dirichlet"], 4)
    d = a.dense()
    for i in k["active"]:
        for j in k["active"]:
            assert d[i, j] == d[j, i]
    assert O.detect_spd(a)


def test_network_rejects_bad_edges():
    with pytest.raises(ValueError):
        _net([(0, 0, 1.0)], {"0": 1.0}, 2)
    with pytest.raises(ValueError):
        _net([(0, 1, -1.0)], {"0": 1.0}, 2)


def test_jacobi_tiny_diagonal(known):
    k = known["jacobi_tiny_diagonal"]
    inv = O.jacobi_inv_diag(csr(np.diag(k["diag"])))
    assert inv[0] == k["inv_diag"][0] and abs(inv[1] - k["inv_diag"][1]) < 1e-12


def test_series_network_flow(known):
    k = known["series_network"]
    g = [1 / r for r in k["resistances"]]
    a, rhs = _net([(0, 1, g[0]), (1, 2, g[1])], {"0": k["p_in"], "2": k["p_out"]}, 3)
    x, path, _ = O.network_cascade(a, rhs, tol=1e-12)
    q1 = g[0] * (x[0] - x[1])
    q2 = g[1] * (x[1] - x[2])
    assert abs(q1 - k["flow"]) <= k["max_relative"] * k["flow"]
    assert abs(q2 - k["flow"]) <= k["max_relative"] * k["flow"]


def test_network_cascade_large_uses_krylov():
    # n > 256 takes the Jacobi-CG branch and meets the equilibrated residual target
    n = 400
    edges = [(i, i + 1, 1.0 + (i % 7)) for i in range(n - 1)] + [(i, i + 20, 0.5) for i in range(0, n - 20, 13)]
    a, rhs = _net(edges, {"0": 1000.0, str(n - 1): 0.0}, n)
    x, path, it = O.network_cascade(a, rhs, tol=1e-10, max_it=5000)
    assert path in (1, 2)
    d = a.dense()
    assert np.linalg.norm(d @ x - rhs) <= 1e-8 * np.linalg.norm(rhs)


# ---- rectangular duct (rectangular.rs:268-300) ------------------------------------------------
def test_rectangular_poiseuille(known):
    k = known["rectangular_duct"]
    po = O.poiseuille_number(1.0)
    assert abs(po - k["square_po"]) <= k["square_po_max_relative"] * k["square_po"]
    # closed form of the fit at eps = 1: 96 / (4 (1 - 192/pi^5 tanh(pi/2)))
    assert abs(po - 96 / (4 * (1 - 192 / math.pi ** 5 * math.tanh(math.pi / 2)))) < 1e-12
    assert O.poiseuille_number(0.0) == 96.0
    mu = 1e-3
    assert abs(O.rect_resistance(1e-3, 5e-4, 0.01, mu) - O.rect_resistance(5e-4, 1e-3, 0.01, mu)) <= 1e-10 * \
        O.rect_resistance(1e-3, 5e-4, 0.01, mu)
    # R = Po mu L / (2 A Dh^2) with Dh = 2wh/(w+h)
    w, h, L = k["dh_w"], k["dh_h"], 0.01
    dh = 2 * w * h / (w + h)
    assert abs(dh - k["dh"]) < 1e-15
    po2 = O.poiseuille_number(h / w)
    assert abs(O.rect_resistance(w, h, L, mu) - po2 * mu * L / (2 * w * h * dh * dh)) <= 1e-12 * \
        O.rect_resistance(w, h, L, mu)


# ---- assembly -----------------------------------------------------------------------------------
def test_pressure_correction_dirichlet_adds_one_over_h2():
    # correction.rs:300-331: the Dirichlet side adds 1/h^2 (not 2/h^2)
    a = O.assemble_pressure_correction(4, 4, 0.25, 0.25, (0, 1, 0, 0)).dense()
    h2 = 16.0
    # cell (i=3, j=1): east boundary (Dirichlet) + west, south, north neighbours
    row = 3 * 4 + 1
    assert a[row, row] == 4 * h2
    # cell (i=0, j=1): west wall Neumann: 3 neighbours
    assert a[1, 1] == 3 * h2
    assert np.allclose(a, a.T)


def test_varcoef_operator_is_symmetric_and_dirichlet_consistent():
    rng = np.random.default_rng(4)
    k = np.exp(rng.uniform(-2, 2, 5 * 4 * 3))
    a, rhs = O.assemble_varcoef(5, 4, 3, 9.0, k, 1 << 4, np.zeros(6))
    d = a.dense()
    assert np.array_equal(d, d.T)
    # constant field: the residual of x = 1 vanishes except on the Dirichlet face
    r = d @ np.ones(60)
    for z in range(3):
        for y in range(4):
            for x in range(5):
                p = (z * 4 + y) * 5 + x
                expect = k[p] * 9.0 if z == 0 else 0.0
                assert abs(r[p] - expect) <= 1e-9 * max(1.0, expect)


def test_momentum_rows_diagonally_dominant():
    rng = np.random.default_rng(9)
    nx = ny = 12
    u = rng.uniform(-1, 1, nx * ny)
    v = rng.uniform(-1, 1, nx * ny)
    a = O.assemble_momentum_u(nx, ny, 1 / 11, 1 / 11, 1e-2, 1e-3, 1.0, u, v).dense()
    for i in range(nx * ny):
        off = np.sum(np.abs(a[i])) - abs(a[i, i])
        assert a[i, i] >= off - 1e-12
