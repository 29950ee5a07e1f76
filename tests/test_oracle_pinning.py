"""Pins the oracle functions that the reference's own tests leave open:
three-phase aggregation (the aggregation of the 128^3 benchmark hierarchy),
the momentum upwind assembly and the variable-coefficient face rule.

Each check is chosen so that a plausible mistake fails it: hand-worked
answers on tiny inputs, properties that follow from the definition of each
phase (checked on the output, not recomputed), and rows written out by hand
from the reference's formulas.  The last test holds the product's
three-phase hierarchy to the oracle bit for bit.  CPU only."""
import numpy as np
import pytest

import cfdrs_amd as ca
from cfdrs_amd import problems as P
from conftest import grid_laplacian, tridiag
from oracle import oracle as O
from test_host_library import _compare_hierarchy


def csr(d):
    return O.Csr.from_dense(np.asarray(d, float))


def strong_rows(a: O.Csr, theta: float):
    """Strong neighbourhoods S_i in column order, from the oracle's strength
    matrix (itself pinned by coarsening/mod.rs:108-122 and brute force)."""
    rp, ci, _ = O.strength(a, theta).arrays()
    return [ci[rp[i]:rp[i + 1]].tolist() for i in range(len(rp) - 1)]


# ---- three-phase aggregation: hand-worked answers -------------------------------------------------
@pytest.mark.parametrize("n,f2c,roots", [
    # 1D Laplacian, every off-diagonal strong.  Phase 1 visits 0 ({0,1}), rejects 2 (1 taken),
    # takes 3 ({2,3,4}), rejects 5, takes 6; phase 2 finds nothing left for n = 7, 8.
    (7, [0, 0, 1, 1, 1, 2, 2], [0, 3, 6]),
    (8, [0, 0, 1, 1, 1, 2, 2, 2], [0, 3, 6]),
    # n = 9: 8's only neighbour 7 is taken, so 8 is left to phase 2 and joins 7's aggregate
    (9, [0, 0, 1, 1, 1, 2, 2, 2, 2], [0, 3, 6]),
    # n = 10: 8 is untouched when 9 is visited, so {9, 8} is a phase-1 aggregate
    (10, [0, 0, 1, 1, 1, 2, 2, 2, 3, 3], [0, 3, 6, 9]),
])
def test_three_phase_known_answer_1d(n, f2c, roots):
    cp, got = O.coarsen(csr(tridiag(n)), "three_phase", 0.25)
    assert got.tolist() == f2c and cp.tolist() == roots


def test_three_phase_known_answer_4x4_grid():
    # Node r*4 + c of the 4x4 5-point grid.  Phase 1: 0 -> {0,1,4}; 2 rejected (1 taken);
    # 3 -> {3,2,7}; 5, 6, 8 rejected; 9 -> {9,5,8,10,13}; 11, 12, 14 rejected; 15 -> {15,11,14}.
    # Phase 2: 6's strong neighbours in column order are 2, 5, 7, 10 -> 2's aggregate (1);
    # 12's are 8, 13 -> 8's aggregate (2).
    cp, f2c = O.coarsen(csr(grid_laplacian(4)), "three_phase", 0.25)
    assert f2c.tolist() == [0, 0, 1, 1, 0, 2, 1, 1, 2, 2, 2, 3, 2, 2, 3, 3]
    assert cp.tolist() == [0, 3, 9, 15]


def test_three_phase_weak_links_split_aggregates():
    # A chain whose (1,2) and (4,5) couplings are weak (0.1 < 0.25 x 1): S_2 = {3}, S_4 = {3},
    # S_5 = {6}.  Phase 1: 0 -> {0,1}; 2 -> {2,3} (its only strong neighbour is 3); 4 rejected
    # (3 taken); 5 -> {5,6}.  Phase 2: 4 joins 3's aggregate.
    d = tridiag(7)
    for i, j in ((1, 2), (4, 5)):
        d[i, j] = d[j, i] = -0.1
    cp, f2c = O.coarsen(csr(d), "three_phase", 0.25)
    assert f2c.tolist() == [0, 0, 1, 1, 1, 2, 2] and cp.tolist() == [0, 2, 5]
    # theta = 0.05 keeps the weak links: 0 -> {0,1}, 2 rejected, 3 -> {2,3,4}, 5 rejected,
    # 6 -> {5,6}
    cp, f2c = O.coarsen(csr(d), "three_phase", 0.05)
    assert f2c.tolist() == [0, 0, 1, 1, 1, 2, 2] and cp.tolist() == [0, 3, 6]


# ---- three-phase aggregation: properties of each phase --------------------------------------------
def _three_phase_cases():
    rng = np.random.default_rng(11)
    out = [("grid10", csr(grid_laplacian(10)), 0.25)]
    for t in range(6):
        m = 40 + 10 * t
        d = -rng.uniform(0.0, 1.0, (m, m)) * (rng.random((m, m)) < 0.08)
        d = d + d.T
        np.fill_diagonal(d, 0.0)
        np.fill_diagonal(d, -d.sum(axis=1) + 0.5)
        out.append((f"random{t}", csr(d), (0.1, 0.25, 0.5)[t % 3]))
    n = 6
    k = P.config2_k(n)
    rp, ci, v, _ = P.varcoef(n, n, n, float(n * n), k, 1 << 4, np.zeros(6))
    out.append(("varcoef6", O.Csr.from_arrays(rp, ci, v, n ** 3), 0.25))
    return out


@pytest.mark.parametrize("name,a,theta", _three_phase_cases(), ids=lambda x: x if isinstance(x, str) else "")
def test_three_phase_phase_properties(name, a, theta):
    n = a.shape[0]
    S = strong_rows(a, theta)
    cp, f2c = O.coarsen(a, "three_phase", theta)
    nc = len(cp)
    assert (f2c >= 0).all() and f2c.max() == nc - 1
    assert all(f2c[r] == idx for idx, r in enumerate(cp))
    # roots are visited in natural order within each phase
    core_of = -np.ones(n, np.int64)  # aggregate whose phase-1 core holds the node
    n1 = 0
    for idx, r in enumerate(cp):
        core = [r] + S[r]
        if any(core_of[j] >= 0 for j in core):
            break  # not a phase-1 aggregate: its neighbourhood was already touched
        core_of[core] = idx
        n1 += 1
    assert list(cp[:n1]) == sorted(cp[:n1])
    # phase 1: each core {root} + S_root lies in its aggregate; cores are pairwise disjoint
    for idx in range(n1):
        for j in [cp[idx]] + S[cp[idx]]:
            assert f2c[j] == idx
    # phase 1 is maximal in visit order: a node that is not a root had, at its visit, itself or a
    # strong neighbour inside the core of an earlier root
    roots1 = set(cp[:n1].tolist())
    for i in range(n):
        if i in roots1:
            continue
        touched = [j for j in [i] + S[i] if core_of[j] >= 0 and cp[core_of[j]] < i]
        assert touched, (name, i)
    # phase 2: a node outside every core joins the aggregate of its first strong neighbour (column
    # order) that lies in a core
    for i in range(n):
        if core_of[i] >= 0:
            continue
        first = next(j for j in S[i] if core_of[j] >= 0)
        assert f2c[i] == core_of[first], (name, i)
    # so nothing is left for phase 3: every point was placed by phase 1 or 2
    assert n1 == nc


# ---- three-phase hierarchy: product vs oracle, bit for bit ---------------------------------------
@pytest.mark.parametrize("n,min_coarse", [(12, 1000), (12, 20), (16, 40)])
def test_amg_three_phase_hierarchy_bit_exact(n, min_coarse):
    """AMGConfig.smoothed_aggregation() (the 128^3 benchmark recipe: three-phase aggregates,
    theta 0.25, smoothed prolongator, Jacobi) on the benchmark's coefficient field: every
    level's A, R, P and aggregate map equal the oracle's bit for bit."""
    k = P.config2_k(n)
    rp, ci, v, _ = P.varcoef(n, n, n, float(n * n), k, 1 << 4, np.zeros(6))
    m = ca.CsrMatrix.from_parts(v, ci, rp, n ** 3, n ** 3)
    cfg = ca.AMGConfig.smoothed_aggregation()
    cfg.min_coarse_size = min_coarse
    amg = ca.AlgebraicMultigrid(m, cfg)
    oamg = O.Amg(O.Csr.from_arrays(rp, ci, v, n ** 3), coarsening="three_phase", interp="smoothed",
                 smoother="jacobi", coarse_solve="direct", theta=cfg.strength_threshold,
                 sa_omega=cfg.smoothing_omega, omega=cfg.relaxation_factor, min_coarse=min_coarse)
    assert oamg.num_levels >= 2
    _compare_hierarchy(amg, oamg)


# ---- momentum upwind rows (coefficients.rs:199-277,406-414; solve.rs:288-303) --------------------
@pytest.mark.parametrize("u_c,v_c,v_e", [(2.0, 1.0, 3.0), (-2.0, 1.0, 3.0), (2.0, -1.0, -3.0),
                                         (-2.0, -1.0, -3.0), (0.0, 0.0, 0.0)])
def test_momentum_row_by_hand(u_c, v_c, v_e):
    """The single interior node of a 3x3 grid.  Diffusion a_e = a_w = mu dy/dx, a_n = a_s =
    mu dx/dy; mass fluxes F_x = rho u dy and F_y = rho v_face dx with v_face the mean of the node's
    v and its east neighbour's.  u > 0 adds F_x to a_w, otherwise subtracts it from a_e; v > 0 adds
    F_y to a_s, otherwise subtracts it from a_n (zero velocity takes the else branch, which adds 0).
    a_P = a_e + a_w + a_n + a_s + rho dx dy / dt; off-diagonals -a_w (i-1), -a_e (i+1),
    -a_s (j-1), -a_n (j+1).  All inputs are dyadic so every value is exact."""
    dx, dy, dt, mu, rho = 0.5, 0.25, 0.5, 0.125, 2.0
    u = np.zeros(9)
    v = np.zeros(9)
    u[4], v[4], v[5] = u_c, v_c, v_e
    a = O.assemble_momentum_u(3, 3, dx, dy, dt, mu, rho, u, v).dense()
    ae = aw = mu * dy / dx           # 0.0625
    an = as_ = mu * dx / dy          # 0.25
    fx = rho * u_c * dy
    fy = rho * ((v_c + v_e) / 2) * dx
    if u_c > 0:
        aw += fx
    else:
        ae -= fx
    if (v_c + v_e) / 2 > 0:
        as_ += fy
    else:
        an -= fy
    expect = np.zeros(9)
    expect[4] = ae + aw + an + as_ + rho * dx * dy / dt
    expect[3], expect[5], expect[1], expect[7] = -aw, -ae, -as_, -an
    assert a[4].tolist() == expect.tolist()
    # the numbers themselves for the first case, written out: F_x = 1, F_y = 2
    if (u_c, v_c) == (2.0, 1.0):
        assert a[4].tolist() == [0, -2.25, 0, -1.0625, 4.125, -0.0625, 0, -0.25, 0]
    if (u_c, v_c) == (-2.0, -1.0):
        assert a[4].tolist() == [0, -0.25, 0, -0.0625, 4.125, -1.0625, 0, -2.25, 0]
    # boundary nodes are identity rows
    for b in (0, 1, 2, 3, 5, 6, 7, 8):
        assert a[b].tolist() == [1.0 if c == b else 0.0 for c in range(9)]


def test_momentum_upwind_side_follows_flow_direction():
    """Reversing the velocity field mirrors the convective coefficient: with u -> -u the
    coefficient that carried F_x moves from west to east (and v likewise north/south)."""
    nx = ny = 7
    rng = np.random.default_rng(5)
    u = rng.uniform(0.5, 1.0, nx * ny)
    v = rng.uniform(0.5, 1.0, nx * ny)
    h = 1.0 / 8.0
    a = O.assemble_momentum_u(nx, ny, h, h, 0.25, 0.0, 1.0, u, v).dense()
    b = O.assemble_momentum_u(nx, ny, h, h, 0.25, 0.0, 1.0, -u, -v).dense()
    for j in range(1, ny - 1):
        for i in range(1, nx - 1):
            p = j * nx + i
            # mu = 0: only convection; positive flow couples west and south, reversed east and north
            assert a[p, p - 1] < 0 and a[p, p + 1] == 0 and a[p, p - nx] < 0 and a[p, p + nx] == 0
            assert b[p, p + 1] == a[p, p - 1] and b[p, p - 1] == 0
            assert b[p, p + nx] == a[p, p - nx] and b[p, p - nx] == 0


# ---- variable-coefficient face rule (harmonic mean) ----------------------------------------------
def test_varcoef_two_cells_harmonic_face():
    """Two cells with k = (1, 3), no Dirichlet face: the shared face carries
    2 k1 k2 / (k1 + k2) = 1.5 times inv_h2 (the arithmetic mean would give 2, the geometric
    sqrt(3))."""
    inv_h2 = 4.0
    a, rhs = O.assemble_varcoef(2, 1, 1, inv_h2, np.array([1.0, 3.0]), 0, np.zeros(6))
    assert a.dense().tolist() == [[6.0, -6.0], [-6.0, 6.0]]
    assert rhs.tolist() == [0.0, 0.0]


def test_varcoef_dirichlet_face_uses_cell_coefficient():
    """A Dirichlet face adds k_P inv_h2 to the diagonal and k_P inv_h2 g to the right-hand side
    (the cell's own coefficient, not a face mean): 3 cells in x with k = (1, 3, 5), Dirichlet
    -x (g = 2) and +x (g = -1)."""
    k = np.array([1.0, 3.0, 5.0])
    dv = np.array([2.0, -1.0, 0.0, 0.0, 0.0, 0.0])
    a, rhs = O.assemble_varcoef(3, 1, 1, 1.0, k, 0b11, dv)
    f01, f12 = 1.5, 2 * 3 * 5 / 8  # harmonic faces: 1.5 and 3.75
    assert a.dense().tolist() == [[f01 + 1.0, -f01, 0.0], [-f01, f01 + f12, -f12], [0.0, -f12, f12 + 5.0]]
    assert rhs.tolist() == [2.0, 0.0, -5.0]
