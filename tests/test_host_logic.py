"""Host-side logic that needs no GPU: the exact-Hessian adjoint iteration, config checks."""

import torch

from torchls_amd.implicit import _exact_adjoint


def test_exact_adjoint_converges_with_an_approximate_preconditioner():
    g = torch.Generator().manual_seed(0)
    n = 40
    A = torch.randn(n, n, dtype=torch.float64, generator=g)
    H = A @ A.T + n * torch.eye(n, dtype=torch.float64)
    # a Gauss-Newton-like approximation: H plus a PSD perturbation
    B = torch.randn(n, 5, dtype=torch.float64, generator=g)
    H_gn = H + 3.0 * B @ B.T
    b = torch.randn(n, dtype=torch.float64, generator=g)
    gn_solve = lambda v: torch.linalg.solve(H_gn, v)  # noqa: E731
    w0 = gn_solve(b)
    exact = torch.linalg.solve(H, b)
    assert torch.linalg.vector_norm(w0 - exact) / torch.linalg.vector_norm(exact) > 1e-3
    w = _exact_adjoint(lambda v: H @ v, gn_solve, b, w0, iterations=30)
    assert torch.linalg.vector_norm(w - exact) / torch.linalg.vector_norm(exact) <= 1e-10
    # zero iterations leave the Gauss-Newton solution unchanged
    assert torch.equal(_exact_adjoint(lambda v: H @ v, gn_solve, b, w0, iterations=0), w0)


def test_conjugate_gradient_config_validates_refinement_limit():
    import pytest

    from torchls_amd.solvers import ConjugateGradientConfig

    assert ConjugateGradientConfig().max_refinements is None
    assert ConjugateGradientConfig(max_refinements=0).max_refinements == 0
    for bad in (-1, True, 1.5):
        with pytest.raises(ValueError, match="max_refinements"):
            ConjugateGradientConfig(max_refinements=bad)
