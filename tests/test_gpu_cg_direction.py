"""GPU parity of the CG direction update folded into the product.

On diagonal storage PCG forms p = z + beta p inside the next product
(EPI_DIR), writing the new direction into the other buffer of a pair; every
other storage runs the separate direction kernel.  Both evaluate the update
as one fused multiply-add, so the two paths differ only in the product's own
rounding (diagonal against row-block CSR sums).  Checked here:

* the same matrix stored both ways takes the same number of iterations
  (within 1) and reaches an explicit FP64 relative residual <= 1e-10, with
  Jacobi and AMG preconditioners, FP64 and FP32-inner with FP64 residual
  replacement;
* the FP32-inner solves stop on the FP32 recursive residual and resume after
  the FP64 replacement: the stopped chunk's remaining products are skipped,
  and the solve must continue from the latest direction (a stale one shows
  up as extra iterations or a residual above the bound);
* solutions of the two paths agree to 1e-7 relative (FP64: 1e-9), far
  inside what one iteration more or less would change.
"""
import numpy as np
import pytest

import cfdrs_amd as ca
from cfdrs_amd import problems as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _system(nn):
    n = nn ** 3
    rp, ci, v, _ = P.varcoef(nn, nn, nn, float(nn * nn), P.config2_k(nn), 1 << 4, np.zeros(6))
    b = P.uniform(91, n)
    return rp, ci, v, b, n


@pytest.mark.parametrize("precision", [ca.Precision.FP64, ca.Precision.FP32_REFINED])
@pytest.mark.parametrize("prec", ["jacobi", "amg"])
def test_fused_direction_matches_separate_update(prec, precision):
    nn = 32
    rp, ci, v, b, n = _system(nn)
    orc = O.Csr.from_arrays(rp, ci, v, n)
    out = {}
    for storage in ("auto", "csr_rowblock"):
        a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
        a.set_storage(storage)
        assert (a.storage()["format"] == "dia") == (storage == "auto")
        m = ca.JacobiPreconditioner(a) if prec == "jacobi" else ca.AlgebraicMultigrid(
            a, ca.AMGConfig.smoothed_aggregation())
        cfg = ca.IterativeSolverConfig.new(1e-10).with_max_iterations(3000).with_precision(precision)
        x = np.zeros(n)
        rep = ca.cg_preconditioned(a, b, m, x, cfg)
        assert rep.termination == ca.Termination.Converged, (storage, rep)
        assert np.linalg.norm(b - orc.spmv(x)) <= 1e-10 * np.linalg.norm(b), storage
        if precision == ca.Precision.FP32_REFINED:
            assert rep.outer_iterations >= 1, rep  # at least one FP64 replacement, then resumed
        out[storage] = (rep.iterations, x)
    it_fused, x_fused = out["auto"]
    it_sep, x_sep = out["csr_rowblock"]
    assert abs(it_fused - it_sep) <= 1, (it_fused, it_sep)
    tol = 1e-9 if precision == ca.Precision.FP64 else 1e-7
    assert np.linalg.norm(x_fused - x_sep) <= tol * np.linalg.norm(x_sep)
