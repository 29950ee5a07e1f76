"""CPU tests of the batched-network host side: synthetic graphs, assembly
bit-exact against the oracle's restatement of matrix_assembly.rs, the
rectangular-duct conductance against the oracle and its width derivative
against central differences."""
import numpy as np
import pytest

from cfdrs_amd import _native as N
from cfdrs_amd import network as NW
from oracle import oracle as O


@pytest.fixture(scope="module")
def nets():
    return NW.synth_networks(count=40, seed_base=4000, nmin=50, nmax=600)


def test_synth_deterministic_and_in_range(nets):
    again = NW.synth_networks(count=40, seed_base=4000, nmin=50, nmax=600)
    assert all(np.array_equal(a, b) for a, b in zip(
        (nets.node_offset, nets.edge_offset, nets.edge_i, nets.edge_j, nets.width, nets.length),
        (again.node_offset, again.edge_offset, again.edge_i, again.edge_j, again.width, again.length)))
    sizes = nets.sizes()
    assert sizes.min() >= 50 and sizes.max() <= 600
    assert nets.width.min() >= 50e-6 and nets.width.max() <= 500e-6
    assert nets.length.min() >= 1e-3 and nets.length.max() <= 10e-3


def test_synth_graphs_connected_planar_degrees(nets):
    for g in range(nets.count):
        n = int(nets.sizes()[g])
        e0, e1 = nets.edge_offset[g], nets.edge_offset[g + 1]
        ei, ej = nets.edge_i[e0:e1], nets.edge_j[e0:e1]
        assert np.all(ei < ej) and np.all(ej < n)
        deg = np.bincount(np.concatenate([ei, ej]), minlength=n)
        assert deg.min() >= 1 and deg.max() <= 4
        assert np.mean(deg >= 2) > 0.9
        parent = list(range(n))

        def find(a):
            while parent[a] != a:
                parent[a] = parent[parent[a]]
                a = parent[a]
            return a

        for a, b in zip(ei, ej):
            parent[find(a)] = find(b)
        assert len({find(k) for k in range(n)}) == 1
        assert len(set(zip(ei.tolist(), ej.tolist()))) == ei.size  # no duplicate channels


def test_conductance_matches_oracle_and_derivative(nets):
    g, dg = NW.rect_conductance(nets.width, NW.CHANNEL_HEIGHT, nets.length)
    for e in range(0, nets.width.size, 37):
        r = O.rect_resistance(nets.width[e], NW.CHANNEL_HEIGHT, nets.length[e], NW.WATER_VISCOSITY)
        assert g[e] == 1.0 / r
    w = nets.width[:50]
    L = nets.length[:50]
    for h in (NW.CHANNEL_HEIGHT, 800e-6):  # w >= h and w < h branches
        _, d = NW.rect_conductance(w, h, L)
        eps = 1e-6 * w
        gp, _ = NW.rect_conductance(w + eps, h, L)
        gm, _ = NW.rect_conductance(w - eps, h, L)
        fd = (gp - gm) / (2 * eps)
        assert np.allclose(d, fd, rtol=1e-6, atol=0)


def test_batch_assembly_matches_oracle(nets):
    batch = NW.channel_batch(nets)
    rp, ci, v, b = batch.export()
    g, _ = NW.rect_conductance(nets.width, NW.CHANNEL_HEIGHT, nets.length)
    for s in range(nets.count):
        n0, n1 = nets.node_offset[s], nets.node_offset[s + 1]
        e0, e1 = nets.edge_offset[s], nets.edge_offset[s + 1]
        n = n1 - n0
        isd = np.zeros(n, np.uint8)
        isd[0] = isd[n - 1] = 1
        dv = np.zeros(n)
        dv[0] = NW.INLET_PRESSURE
        a, rhs = O.assemble_network(n, nets.edge_i[e0:e1], nets.edge_j[e0:e1], g[e0:e1], isd, dv)
        orp, oci, ov = a.arrays()
        assert np.array_equal(rp[n0:n1 + 1] - rp[n0], orp)
        assert np.array_equal(ci[rp[n0]:rp[n1]] - n0, oci)
        assert np.array_equal(v[rp[n0]:rp[n1]], ov)
        assert np.array_equal(b[n0:n1], rhs)
        d = a.dense()
        assert np.array_equal(d, d.T)  # symmetric: Dirichlet rows are identity, their columns eliminated


def test_batch_rejects_invalid_networks():
    one = dict(node_offset=[0, 3], edge_offset=[0, 2], is_dirichlet=[1, 0, 1], dirichlet_value=[1.0, 0.0, 0.0])
    NW.NetworkBatch(one["node_offset"], one["edge_offset"], [0, 1], [1, 2], [1.0, 2.0], one["is_dirichlet"],
                    one["dirichlet_value"])
    with pytest.raises(N.CfdlaError, match="Self-loop"):
        NW.NetworkBatch([0, 3], [0, 2], [0, 1], [1, 1], [1.0, 2.0], [1, 0, 1], [1.0, 0, 0])
    with pytest.raises(N.CfdlaError, match="Non-positive"):
        NW.NetworkBatch([0, 3], [0, 2], [0, 1], [1, 2], [1.0, -2.0], [1, 0, 1], [1.0, 0, 0])
    with pytest.raises(N.CfdlaError, match="Dirichlet"):
        NW.NetworkBatch([0, 3], [0, 2], [0, 1], [1, 2], [1.0, 2.0], [0, 0, 0], [0.0, 0, 0])
    with pytest.raises(N.CfdlaError, match="Isolated"):
        NW.NetworkBatch([0, 4], [0, 2], [0, 1], [1, 2], [1.0, 2.0], [1, 0, 1, 0], [1.0, 0, 0, 0])
    with pytest.raises(N.CfdlaError, match="2048"):
        n = 2100
        NW.NetworkBatch([0, n], [0, n - 1], np.arange(n - 1), np.arange(1, n), np.ones(n - 1),
                        np.r_[1, np.zeros(n - 1)].astype(np.uint8), np.zeros(n))


def test_three_node_known_answer(known):
    k = known["network_three_node"]
    b = NW.NetworkBatch([0, 3], [0, 2], [0, 1], [1, 2], [1.0, 2.0], [1, 0, 1], [100.0, 0.0, 0.0],
                        neumann=[0.0, k["neumann_b"], 0.0])
    rp, ci, v, rhs = b.export()
    d = np.zeros((3, 3))
    for i in range(3):
        for q in range(rp[i], rp[i + 1]):
            d[i, ci[q]] += v[q]
    assert d[1, 1] == k["diag_b"] and d[0, 0] == 1.0 and d[2, 2] == 1.0 and d[1, 0] == 0.0
    assert rhs[1] == k["rhs_with_neumann_b"]


def _network(nets, s):
    """Local arrays of network s with the channel-batch boundary conditions:
    node 0 at the inlet pressure, the last node grounded and the outlet."""
    n = int(nets.node_offset[s + 1] - nets.node_offset[s])
    e0, e1 = nets.edge_offset[s], nets.edge_offset[s + 1]
    is_dir = np.zeros(n, np.uint8)
    dval = np.zeros(n)
    is_dir[[0, n - 1]] = 1
    dval[0] = NW.INLET_PRESSURE
    outlet = np.zeros(n)
    outlet[n - 1] = 1.0
    return n, nets.edge_i[e0:e1], nets.edge_j[e0:e1], nets.width[e0:e1].copy(), nets.length[e0:e1], is_dir, dval, outlet


def test_oracle_conductance_complex_step_matches_resistance_and_differences():
    rng = np.random.default_rng(9)
    w = rng.uniform(51e-6, 500e-6, 64)
    L = rng.uniform(1e-3, 10e-3, 64)
    g, dg = O.rect_conductance_complex_step(w, NW.CHANNEL_HEIGHT, L, NW.WATER_VISCOSITY)
    r = np.array([O.rect_resistance(a, NW.CHANNEL_HEIGHT, b, NW.WATER_VISCOSITY) for a, b in zip(w, L)])
    assert np.max(np.abs(g * r - 1.0)) <= 4e-16 * 4
    h = 1e-6 * w
    gp, _ = O.rect_conductance_complex_step(w + h, NW.CHANNEL_HEIGHT, L, NW.WATER_VISCOSITY)
    gm, _ = O.rect_conductance_complex_step(w - h, NW.CHANNEL_HEIGHT, L, NW.WATER_VISCOSITY)
    assert np.max(np.abs((gp - gm) / (2 * h) / dg - 1.0)) <= 1e-7
    # the host library's closed-form width derivative agrees with the complex step
    g_lib, dg_lib = NW.rect_conductance(w, NW.CHANNEL_HEIGHT, L)
    assert np.max(np.abs(g_lib / g - 1.0)) <= 1e-14
    assert np.max(np.abs(dg_lib / dg - 1.0)) <= 1e-12


def test_oracle_network_gradient_matches_central_differences():
    # pins the oracle's adjoint network gradient (assembly, solve, adjoint,
    # edge formula and complex-step dG/dw) by central differences of its own
    # outlet flux along random relative width perturbations
    nets = NW.synth_networks(count=3, seed_base=4200, nmin=60, nmax=160)
    rng = np.random.default_rng(42)
    for s in range(nets.count):
        n, ei, ej, w, L, is_dir, dval, outlet = _network(nets, s)
        J, grad, p = O.network_outlet_gradient(n, ei, ej, w, NW.CHANNEL_HEIGHT, L, NW.WATER_VISCOSITY, is_dir,
                                               dval, outlet)
        # J is the total flow into the outlet: with one inlet and one outlet
        # it equals the flow out of the inlet (mass conservation)
        at_in = (ei == 0) | (ej == 0)
        g, _ = O.rect_conductance_complex_step(w, NW.CHANNEL_HEIGHT, L, NW.WATER_VISCOSITY)
        other = np.where(ei == 0, ej, ei)
        q_in = np.sum(g[at_in] * (p[0] - p[other[at_in]]))
        assert abs(J - q_in) <= 1e-9 * abs(J)
        assert J > 0
        t = 1e-6
        for _ in range(6):
            d = rng.normal(size=w.size) * w
            jp = O.network_outlet_gradient(n, ei, ej, w + t * d, NW.CHANNEL_HEIGHT, L, NW.WATER_VISCOSITY, is_dir,
                                           dval, outlet)[0]
            jm = O.network_outlet_gradient(n, ei, ej, w - t * d, NW.CHANNEL_HEIGHT, L, NW.WATER_VISCOSITY, is_dir,
                                           dval, outlet)[0]
            fd = (jp - jm) / (2 * t)
            assert abs(fd - grad @ d) <= 1e-6 * abs(fd) + 1e-12 * np.abs(grad * d).sum(), (s, fd, grad @ d)


def test_oracle_network_gradient_plausible_mistakes_fail():
    # the same check catches a flipped adjoint sign and a dropped explicit term
    nets = NW.synth_networks(count=1, seed_base=4210, nmin=60, nmax=90)
    n, ei, ej, w, L, is_dir, dval, outlet = _network(nets, 0)
    args = (NW.CHANNEL_HEIGHT, L, NW.WATER_VISCOSITY, is_dir, dval, outlet)
    J, grad, p = O.network_outlet_gradient(n, ei, ej, w, *args)
    g, dg = O.rect_conductance_complex_step(w, NW.CHANNEL_HEIGHT, L, NW.WATER_VISCOSITY)
    m = outlet
    explicit = m[ej] * (p[ei] - p[ej]) + m[ei] * (p[ej] - p[ei])
    d = np.random.default_rng(3).normal(size=w.size) * w
    t = 1e-6
    fd = (O.network_outlet_gradient(n, ei, ej, w + t * d, *args)[0] -
          O.network_outlet_gradient(n, ei, ej, w - t * d, *args)[0]) / (2 * t)
    no_explicit = grad - explicit * dg
    assert abs(no_explicit @ d - fd) > 1e-3 * abs(fd)
    flipped = 2 * explicit * dg - grad  # explicit + (lambda_i - lambda_j)(p_i - p_j), times dG/dw
    assert abs(flipped @ d - fd) > 1e-3 * abs(fd)
