"""GPU parity of the batched network solver against dense FP64 solves of the
oracle-assembled systems, at small size per network, and through
size-independent properties (flow conservation, equilibrated residual) at the
full 4,096-network size.  Bar: equilibrated relative residual <= 1e-10 per
network (the reference's measure, linear_system.rs:188-246) recomputed on the
host, and solution error against the dense solve <= 1e-7 (relative)."""
import numpy as np
import pytest

import cfdrs_amd as ca
from cfdrs_amd import network as NW

pytestmark = pytest.mark.gpu


def eq_residuals(batch, x):
    rp, ci, v, b = batch.export()
    rows = np.repeat(np.arange(batch.n_rows), np.diff(rp))
    ax = np.bincount(rows, weights=v * x[ci], minlength=batch.n_rows)
    rmax = np.maximum.reduceat(np.abs(v), rp[:-1])
    eq = np.where(rmax > np.finfo(float).eps, 1.0 / rmax, 1.0)
    r2 = (eq * (b - ax)) ** 2
    b2 = (eq * b) ** 2
    no = batch.node_offset
    return np.sqrt(np.add.reduceat(r2, no[:-1]) / np.add.reduceat(b2, no[:-1]))


def dense_solutions(batch, nets):
    rp, ci, v, b = batch.export()
    out = np.empty(batch.n_rows)
    for s in range(batch.n_sys):
        n0, n1 = nets.node_offset[s], nets.node_offset[s + 1]
        d = np.zeros((n1 - n0, n1 - n0))
        for i in range(n0, n1):
            for q in range(rp[i], rp[i + 1]):
                d[i - n0, ci[q] - n0] += v[q]
        out[n0:n1] = np.linalg.solve(d, b[n0:n1])
    return out


@pytest.fixture(scope="module")
def small():
    nets = NW.synth_networks(count=48, seed_base=4000, nmin=50, nmax=700)
    return nets, NW.channel_batch(nets)


@pytest.mark.parametrize("precision,reliable", [(ca.Precision.FP64, 0), (ca.Precision.FP32_REFINED, 0),
                                                (ca.Precision.FP32_REFINED, -1)],
                         ids=["fp64", "fp32-reliable", "fp32-restarted"])
@pytest.mark.parametrize("wave_max", [0, 256, 512])
def test_batch_matches_dense(small, precision, reliable, wave_max):
    nets, batch = small
    batch.schedule = {"wave_max_rows": wave_max, "blocks_per_cu": 2}
    cfg = (ca.IterativeSolverConfig.new(1e-10).with_max_iterations(20000).with_precision(precision)
           .with_reliable_interval(reliable))
    x, reps = batch.solve(None, cfg)
    assert all(r.termination == ca.Termination.Converged for r in reps), [r for r in reps if
                                                                          r.termination != ca.Termination.Converged][:3]
    res = eq_residuals(batch, x)
    assert res.max() <= 1e-10
    xs = dense_solutions(batch, nets)
    for s in range(batch.n_sys):
        n0, n1 = nets.node_offset[s], nets.node_offset[s + 1]
        err = np.linalg.norm(x[n0:n1] - xs[n0:n1]) / np.linalg.norm(xs[n0:n1])
        assert err <= 1e-7, (s, err)
    # reported norms agree with the host recomputation; two evaluations of
    # ||D(b - A x)|| / ||D b|| in different summation orders differ by about
    # (row length) * eps * ||D(|b| + |A||x|)|| / ||D b||, i.e. ~1e-15 here
    rep_res = np.array([r.final_residual_norm / max(r.threshold / 1e-10, 1e-300) for r in reps])
    assert np.allclose(rep_res, res, rtol=1e-6, atol=1e-14)


def _dense_plus_synth(nets):
    # the synthetic networks and one 300-node network whose first 160 nodes
    # are all connected (25,000+ entries, rows of up to 160)
    nd = 300
    ei = [(a, b) for a in range(160) for b in range(a + 1, 160)] + [(a, a + 1) for a in range(159, nd - 1)]
    e_i = np.concatenate([nets.edge_i, [a for a, _ in ei]])
    e_j = np.concatenate([nets.edge_j, [b for _, b in ei]])
    no = np.concatenate([nets.node_offset, [nets.node_offset[-1] + nd]])
    eo = np.concatenate([nets.edge_offset, [nets.edge_offset[-1] + len(ei)]])
    g = np.concatenate([NW.rect_conductance(nets.width, NW.CHANNEL_HEIGHT, nets.length)[0],
                        1e-12 * (1.0 + np.random.default_rng(7).random(len(ei)))])
    is_dir = np.zeros(no[-1], np.uint8)
    dval = np.zeros(no[-1])
    is_dir[no[:-1]] = 1
    dval[no[:-1]] = NW.INLET_PRESSURE
    is_dir[no[1:] - 1] = 1
    return NW.NetworkBatch(no, eo, e_i, e_j, g, is_dir, dval)


@pytest.mark.parametrize("precision,reliable", [(ca.Precision.FP64, 0), (ca.Precision.FP32_REFINED, 0),
                                                (ca.Precision.FP32_REFINED, -1)],
                         ids=["fp64", "fp32-reliable", "fp32-restarted"])
def test_schedule_does_not_change_results(small, precision, reliable):
    # the persistent grid's size only changes which block picks up which
    # system: every team's arithmetic is the same, so the solutions,
    # iteration counts and reported norms are bit-identical between one and
    # two blocks per CU, with a dense system (long rows) in the batch
    nets, _ = small
    batch = _dense_plus_synth(nets)
    cfg = (ca.IterativeSolverConfig.new(1e-10).with_max_iterations(20000).with_precision(precision)
           .with_reliable_interval(reliable))
    out = {}
    for bpc in (1, 2):
        batch.schedule = {"wave_max_rows": 0, "blocks_per_cu": bpc}
        out[bpc] = batch.solve(None, cfg)
    (x1, r1), (x2, r2) = out[1], out[2]
    assert np.array_equal(x1.view(np.uint64), x2.view(np.uint64))
    assert [r.iterations for r in r1] == [r.iterations for r in r2]
    assert [r.final_residual_norm for r in r1] == [r.final_residual_norm for r in r2]
    assert all(r.termination == ca.Termination.Converged for r in r1)
    assert eq_residuals(batch, x1).max() <= 1e-10


def test_batch_reliable_updates_keep_fp64_iteration_counts(small):
    # FP32 recurrence with FP64 residual replacement and adaptive delta: the
    # iteration counts stay near the FP64 ones (restarted refinement roughly
    # doubles them on these ill-conditioned Laplacians)
    _, batch = small
    batch.schedule = {"wave_max_rows": 256, "blocks_per_cu": 2}
    base = ca.IterativeSolverConfig.new(1e-10).with_max_iterations(20000)
    _, r64 = batch.solve(None, base)
    _, r32 = batch.solve(None, ca.IterativeSolverConfig.new(1e-10).with_max_iterations(20000)
                         .with_precision(ca.Precision.FP32_REFINED))
    i64 = np.array([r.iterations for r in r64])
    i32 = np.array([r.iterations for r in r32])
    assert np.all(i32 <= 2 * i64 + 20), np.max(i32 / i64)
    assert i32.mean() <= 1.5 * i64.mean(), (i32.mean(), i64.mean())
    assert all(r.outer_iterations >= 1 for r in r32)


def test_batch_torch_device_and_warm_start(small):
    import torch

    nets, batch = small
    x = torch.zeros(batch.n_rows, dtype=torch.float64, device="cuda")
    cfg = ca.IterativeSolverConfig.new(1e-10).with_max_iterations(20000)
    _, reps = batch.solve(x, cfg)
    assert all(r.termination == ca.Termination.Converged for r in reps)
    _, reps2 = batch.solve(x, cfg)  # already converged: initial residual
    assert all(r.termination == ca.Termination.InitialResidual and r.iterations == 0 for r in reps2)


def test_autotune_picks_a_candidate(small, tmp_path):
    _, batch = small
    cfg = ca.IterativeSolverConfig.new(1e-10).with_max_iterations(20000)
    cands = [(0, 2), (256, 2), (512, 1)]
    entry = batch.autotune(cfg, candidates=cands, cache_path=str(tmp_path / "tune.json"))
    assert (entry["best"]["wave_max_rows"], entry["best"]["blocks_per_cu"]) in cands
    assert batch.schedule == entry["best"]
    again = batch.autotune(cfg, candidates=cands, cache_path=str(tmp_path / "tune.json"))
    assert again == entry  # served from the JSON cache


def test_full_size_batch_properties():
    nets = NW.synth_networks()  # 4096 networks, 50..2000 nodes
    batch = NW.channel_batch(nets)
    cfg = ca.IterativeSolverConfig.new(1e-10).with_max_iterations(50000)
    x, reps = batch.solve(None, cfg)
    assert len(reps) == 4096
    assert all(r.termination == ca.Termination.Converged for r in reps)
    assert eq_residuals(batch, x).max() <= 1e-10
    # flow conservation at free nodes: sum of G (p_j - p_i) vanishes
    g = batch.conductance
    sys_of_edge = np.repeat(np.arange(nets.count), np.diff(nets.edge_offset))
    base = nets.node_offset[:-1][sys_of_edge]
    i, j = nets.edge_i + base, nets.edge_j + base
    q = g * (x[i] - x[j])
    net = np.bincount(i, weights=-q, minlength=batch.n_rows) + np.bincount(j, weights=q, minlength=batch.n_rows)
    free = batch.is_dirichlet == 0
    # scale: the node's total conductance times the pressure range (the inlet
    # pressure); a residual of 1e-10 in the row-equilibrated norm leaves
    # |net| / scale near 1e-10, a dropped edge leaves G * dp / scale ~ 1e-3
    scale = (np.bincount(i, weights=g, minlength=batch.n_rows) +
             np.bincount(j, weights=g, minlength=batch.n_rows)) * NW.INLET_PRESSURE
    assert np.max(np.abs(net[free]) / np.maximum(scale[free], 1e-300)) <= 1e-8
    # pressures bounded by the boundary values (discrete maximum principle),
    # up to the 1e-7 relative solution error a 1e-10 residual allows
    assert x.min() >= -1e-7 * NW.INLET_PRESSURE and x.max() <= NW.INLET_PRESSURE * (1 + 1e-7)


def test_width_gradient_matches_finite_differences():
    import torch

    nets = NW.synth_networks(count=6, seed_base=4100, nmin=60, nmax=300)
    batch = NW.channel_batch(nets)
    cfg = ca.IterativeSolverConfig.new(1e-12).with_max_iterations(50000)
    g, dg_dw = NW.rect_conductance(nets.width, NW.CHANNEL_HEIGHT, nets.length)

    def J(width):
        gg, _ = NW.rect_conductance(width, NW.CHANNEL_HEIGHT, nets.length)
        batch.update_conductance(gg)
        x, _ = batch.solve(None, cfg)
        return NW.outlet_flux(nets, gg, x), x

    J0, p = J(nets.width)
    # adjoint: A lambda = dJ/dp, then the edge-gradient kernel
    sys_of_edge = np.repeat(np.arange(nets.count), np.diff(nets.edge_offset))
    base = nets.node_offset[:-1][sys_of_edge]
    outlet = (nets.node_offset[1:] - 1)[sys_of_edge] - base
    at_i, at_j = nets.edge_i == outlet, nets.edge_j == outlet
    other = np.where(at_i, nets.edge_j, nets.edge_i) + base
    touch = at_i | at_j
    dJdp = np.zeros(batch.n_rows)
    np.add.at(dJdp, other[touch], g[touch])
    dJdp[batch.is_dirichlet == 1] = 0.0
    explicit = np.where(touch, p[other] - p[np.where(at_i, nets.edge_i, nets.edge_j) + base], 0.0)
    batch.update_conductance(g)
    lam = torch.zeros(batch.n_rows, dtype=torch.float64, device="cuda")
    _, reps = batch.solve(lam, cfg, rhs=torch.from_numpy(dJdp).cuda())
    assert all(r.termination in (ca.Termination.Converged, ca.Termination.InitialResidual) for r in reps)
    grad = batch.edge_gradient(torch.from_numpy(p).cuda(), lam, torch.from_numpy(explicit).cuda(),
                               torch.from_numpy(dg_dw).cuda()).cpu().numpy()
    rng = np.random.default_rng(1)
    for e in rng.choice(nets.width.size, size=12, replace=False):
        h = 1e-4 * nets.width[e]
        wp, wm = nets.width.copy(), nets.width.copy()
        wp[e] += h
        wm[e] -= h
        s = sys_of_edge[e]
        fd = (J(wp)[0][s] - J(wm)[0][s]) / (2 * h)
        assert abs(grad[e] - fd) <= 1e-5 * max(abs(fd), 1e-3 * np.max(np.abs(grad))), (e, grad[e], fd)


def test_device_reassembly_matches_host_bit_for_bit(small):
    # new conductances formed on the GPU: values, right-hand side and the
    # solve equal those of the host assembly from the same conductances
    import torch

    nets, _ = small
    rng = np.random.default_rng(17)
    g0 = NW.rect_conductance(nets.width, NW.CHANNEL_HEIGHT, nets.length)[0]
    g1 = g0 * rng.uniform(0.5, 2.0, g0.size)
    host = NW.channel_batch(nets)
    host.update_conductance(g1)
    dev = NW.channel_batch(nets)
    dev.update_conductance_device(torch.from_numpy(g1).cuda())
    rp_h, ci_h, v_h, b_h = host.export()
    rp_d, ci_d, v_d, b_d = dev.export()
    assert np.array_equal(rp_h, rp_d) and np.array_equal(ci_h, ci_d)
    assert np.array_equal(v_h, v_d)
    assert np.array_equal(b_h, b_d)
    cfg = ca.IterativeSolverConfig.new(1e-10).with_max_iterations(20000)
    xh, _ = host.solve(None, cfg)
    xd, reps = dev.solve(None, cfg)
    assert all(r.termination == ca.Termination.Converged for r in reps)
    assert np.array_equal(xh, xd)  # same values, same kernel, same schedule
    assert np.array_equal(dev.conductance, g1)


def test_device_reassembly_rejects_bad_conductance(small):
    import torch

    nets, batch = small
    before = batch.export()
    g = torch.from_numpy(batch.conductance.copy()).cuda()
    g[3] = -1.0
    with pytest.raises(ca.CfdlaError):
        batch.update_conductance_device(g)
    g[3] = float("nan")
    with pytest.raises(ca.CfdlaError):
        batch.update_conductance_device(g)
    after = batch.export()
    assert all(np.array_equal(a, b) for a, b in zip(before, after))


def test_device_conductance_matches_host():
    # Po(eps) fit and dG/dw on the device: same formulas, device tanh, so
    # agreement to a few ulp (relative 1e-15); widths 50-500 um at h = 50 um
    import torch

    nets = NW.synth_networks(count=16, seed_base=4100, nmin=50, nmax=400)
    gh, dh = NW.rect_conductance(nets.width, NW.CHANNEL_HEIGHT, nets.length)
    gd, dd = NW.rect_conductance_device(torch.from_numpy(nets.width).cuda(), NW.CHANNEL_HEIGHT,
                                        torch.from_numpy(nets.length).cuda())
    assert np.allclose(gd.cpu().numpy(), gh, rtol=1e-15, atol=0)
    assert np.allclose(dd.cpu().numpy(), dh, rtol=1e-14, atol=0)


def test_design_loop_on_device(small):
    # widths -> conductances -> assembly -> solve, all on the GPU, equals the
    # host pipeline from the same widths
    import torch

    nets, _ = small
    w = nets.width * 1.3
    dev = NW.channel_batch(nets)
    gd, _ = NW.rect_conductance_device(torch.from_numpy(w).cuda(), NW.CHANNEL_HEIGHT,
                                       torch.from_numpy(nets.length).cuda())
    dev.update_conductance_device(gd)
    x = torch.zeros(dev.n_rows, dtype=torch.float64, device="cuda")
    cfg = ca.IterativeSolverConfig.new(1e-10).with_max_iterations(20000)
    _, reps = dev.solve(x, cfg)
    assert all(r.termination == ca.Termination.Converged for r in reps)
    host = NW.channel_batch(nets)
    host.update_conductance(NW.rect_conductance(w, NW.CHANNEL_HEIGHT, nets.length)[0])
    xh, _ = host.solve(None, cfg)
    assert np.max(np.abs(x.cpu().numpy() - xh)) <= 1e-9 * NW.INLET_PRESSURE
