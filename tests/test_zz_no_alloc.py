"""Allocation-free batched network solve (cfdla_batch_solve_device_noalloc)
against the allocating entry point it mirrors (cfdla_batch_solve_device),
called through the C ABI on two batches in a row, the second one smaller,
in both precisions.  Each batch's first cfdla_batch_solve_device puts it on
the device; the allocation-free call then runs on the same inputs.

Bar: bit-exact.  Both calls launch the same kernels on the same data with
the same schedule; only the work-queue counter and the report staging
differ in where they live, and each system's arithmetic is independent of
which team takes it.
"""
import ctypes as C

import numpy as np
import pytest

import cfdrs_amd as ca
from cfdrs_amd import _native as N
from cfdrs_amd import network as NW
from cfdrs_amd.sparse import bind_torch_stream

pytestmark = pytest.mark.gpu


def _run(fn, batch, cfg, x):
    reps = (N.Report * batch.n_sys)()
    bind_torch_stream()
    N.call(fn, batch.handle, C.byref(cfg), None, x.data_ptr(), reps)
    return [(reps[k].iterations, reps[k].outer_iterations, reps[k].initial_residual_norm,
             reps[k].final_residual_norm, reps[k].threshold, reps[k].termination) for k in range(batch.n_sys)]


@pytest.mark.parametrize("precision", [ca.Precision.FP64, ca.Precision.FP32_REFINED])
def test_noalloc_batch_solve_matches_allocating_solve(precision):
    import torch

    cfg = (ca.IterativeSolverConfig.new(1e-10).with_max_iterations(20000).with_precision(precision)
           .native(N.KIND_CG))
    for count, seed, nmax in ((40, 4300, 900), (13, 4400, 300)):  # the second batch is smaller
        nets = NW.synth_networks(count=count, seed_base=seed, nmin=50, nmax=nmax)
        batch = NW.channel_batch(nets)
        x_ref = torch.zeros(batch.n_rows, dtype=torch.float64, device="cuda:0")
        rep_ref = _run("cfdla_batch_solve_device", batch, cfg, x_ref)
        x_na = torch.full_like(x_ref, 0.0)
        rep_na = _run("cfdla_batch_solve_device_noalloc", batch, cfg, x_na)
        assert torch.equal(x_na, x_ref), (count, precision)
        assert rep_na == rep_ref
        assert all(r[5] == N.TERM_CONVERGED for r in rep_na)
        # a second allocation-free call on the same batch starts its queue from zero again
        x_na2 = torch.zeros_like(x_ref)
        assert _run("cfdla_batch_solve_device_noalloc", batch, cfg, x_na2) == rep_ref
        assert torch.equal(x_na2, x_ref)


def test_noalloc_batch_solve_rejects_a_batch_not_on_the_device():
    nets = NW.synth_networks(count=3, seed_base=4500, nmin=50, nmax=80)
    batch = NW.channel_batch(nets)
    import torch

    x = torch.zeros(batch.n_rows, dtype=torch.float64, device="cuda:0")
    cfg = ca.IterativeSolverConfig.new(1e-10).native(N.KIND_CG)
    reps = (N.Report * batch.n_sys)()
    rc = N.lib().cfdla_batch_solve_device_noalloc(batch.handle, C.byref(cfg), None, C.c_void_p(x.data_ptr()), reps)
    assert rc == N.ERR_CONFIG
    assert b"resident" in N.lib().cfdla_last_error()
    assert torch.count_nonzero(x) == 0
