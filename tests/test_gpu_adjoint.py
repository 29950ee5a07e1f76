"""GPU: dJ/dg of the 256^2 variable-coefficient problem (GRF log-coefficient,
correlation length 16 cells, sigma 1) from the adjoint AMG-PCG solve and the
coefficient-gradient kernel, checked against the CPU oracle's adjoint
gradient (relative 1e-6, both solves at >= 1e-10) and against central finite
differences of GPU forward solves on 32 random directions (relative 1e-5
plus a floor of 1e-8 |grad| |d| for nearly cancelling directions),
and through torch.autograd."""
import numpy as np
import pytest

from cfdrs_amd import adjoint as ADJ
from cfdrs_amd import problems as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NX = NY = 256


@pytest.fixture(scope="module")
def setup():
    g = P.grf(NX, NY, 16.0, seed=P.SEED_C5)
    prob = ADJ.FluxProblem(NX, NY, tol=1e-12)
    J, grad, p, lam = ADJ.flux_and_gradient(prob, g)
    return g, prob, J, grad.cpu().numpy(), p.cpu().numpy()


def test_gradient_matches_oracle_adjoint(setup):
    g, prob, J, grad, p = setup
    k = np.exp(g)
    a, rhs = O.assemble_varcoef(NX, NY, 1, prob.inv_h2, k, ADJ.DIR_MASK, np.array(ADJ.DIR_VAL))
    po, _ = O.solve(a, rhs, kind="cg", prec="jacobi", tol_rel=1e-12, max_it=200000)
    assert abs(prob.objective(k, po) - J) <= 1e-8 * abs(J)
    lam, _ = O.solve(a, prob.dJ_dp(k), kind="cg", prec="jacobi", tol_rel=1e-12, max_it=200000)
    go = O.coef_grad_2d(NX, NY, prob.inv_h2, ADJ.DIR_MASK, np.array(ADJ.DIR_VAL[:4]), k, po, lam,
                        prob.dJ_dk_explicit(po))
    assert np.linalg.norm(grad - go) <= 1e-6 * np.linalg.norm(go)


def test_gradient_matches_central_differences(setup):
    g, prob, J, grad, _ = setup
    rng = np.random.default_rng(5000)
    h = 1e-4
    for d_i in range(32):
        d = rng.normal(size=NX * NY)
        jp = ADJ.flux_and_gradient(prob, g + h * d)[0]
        jm = ADJ.flux_and_gradient(prob, g - h * d)[0]
        fd = (jp - jm) / (2 * h)
        # relative 1e-5, with a floor of 1e-8 |grad| |d| (the Cauchy-Schwarz
        # scale of grad . d) for directions where the derivative nearly
        # cancels: the O(h^2) truncation error does not shrink with it
        tol = 1e-5 * abs(fd) + 1e-8 * np.linalg.norm(grad) * np.linalg.norm(d)
        assert abs(fd - grad @ d) <= tol, (d_i, fd, grad @ d)


def test_torch_autograd_function(setup):
    import torch

    g, prob, J, grad, _ = setup
    gt = torch.tensor(g, dtype=torch.float64, device="cuda", requires_grad=True)
    out = ADJ.outlet_flux(gt, prob)
    (3.0 * out).backward()
    assert abs(out.item() - J) <= 1e-12 * abs(J)
    assert np.allclose(gt.grad.cpu().numpy(), 3.0 * grad, rtol=1e-12, atol=0)
