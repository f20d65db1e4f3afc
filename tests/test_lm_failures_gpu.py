"""LM termination and failure paths on the HIP backend (solvers.py:343-345,
386-397, 595-596, 690-693): each reason string of the reference, forced by a
constructed problem whose expected outcome is derived in the test."""

import math

import numpy as np
import pytest
import torch

from torchls_amd import Problem, TerminationConfig, TrustRegionConfig

pytestmark = pytest.mark.gpu


def sqrt_residual(x):
    return torch.sqrt(x)


def atan_residual(x):
    return torch.atan(x)


def linear_residual(x, target):
    return x - target


def sum_residual(a, b, t):
    return a[:3] + a[3:] + b - t


def two_type_problem(device):
    """Residual a[:3] + a[3:] + b - t (3 rows; 3x6 and 3x3 blocks, the shapes of the
    HIP Schur kernels); a[2] and b[3] are never observed, and J_a^T J_a = [I I; I I]
    is singular for every a."""
    P = Problem({"a": torch.ones(3, 6, dtype=torch.float64, device=device),
                 "b": torch.ones(4, 3, dtype=torch.float64, device=device)})
    P.add_residual(sum_residual, ("a", torch.tensor([0, 1, 0])), ("b", torch.tensor([0, 1, 2])),
                   torch.zeros(3, 3, dtype=torch.float64, device=device))
    return P


def test_nonfinite_linearization(device):
    """sqrt at 0: finite residual (cost 0 is finite), infinite derivative."""
    print("lm nonfinite_linearization", flush=True)
    P = Problem({"x": torch.zeros(4, 1, dtype=torch.float64, device=device)})
    P.add_residual(sqrt_residual, ("x", torch.arange(4)))
    _, s = P.solve(return_summary=True)
    assert s.termination_reason == "nonfinite_linearization"
    assert (s.iterations, s.accepted_steps, s.rejected_steps) == (1, 0, 0)


def test_initial_nonfinite_raises(device):
    P = Problem({"x": torch.full((2, 1), -1.0, dtype=torch.float64, device=device)})
    P.add_residual(sqrt_residual, ("x", torch.arange(2)))
    with pytest.raises(ValueError, match="Initial residuals are nonfinite"):
        P.solve()


def test_no_acceptable_step(device):
    """atan(x) from x = 2: the Gauss-Newton step of the linearization (slope 1/5)
    jumps past -3, where the cost is larger; with two trials both steps are
    rejected (checked here in float64 numpy) and the solve stops."""
    print("lm no_acceptable_step", flush=True)
    x0 = 2.0
    P = Problem({"x": torch.tensor([[x0]], dtype=torch.float64, device=device)})
    P.add_residual(atan_residual, ("x", torch.arange(1)))
    tr = TrustRegionConfig(max_trials=2)
    sol, s = P.solve(return_summary=True, trust_region=tr)
    j, r = 1.0 / (1.0 + x0 * x0), math.atan(x0)
    lams = [tr.lambda_initial, tr.lambda_initial * tr.lambda_factor]
    for lam in lams:  # the expected rejections
        step = -(j * r) / (j * j + lam)
        assert math.atan(x0 + step) ** 2 > r * r
    assert s.termination_reason == "no_acceptable_step"
    assert (s.iterations, s.accepted_steps, s.rejected_steps) == (1, 0, 2)
    assert s.lambda_history == pytest.approx(lams, rel=1e-15)
    assert float(sol["x"][0, 0]) == x0


def test_time_limit(device):
    """A time limit already spent stops before the first iteration."""
    P = Problem({"x": torch.ones(3, 1, dtype=torch.float64, device=device)})
    P.add_residual(linear_residual, ("x", torch.arange(3)), torch.zeros(3, 1, dtype=torch.float64, device=device))
    _, s = P.solve(return_summary=True, termination=TerminationConfig(max_time_seconds=1e-12))
    assert s.termination_reason == "time_limit" and s.iterations == 0


@pytest.mark.parametrize("solver", ["conjugate_gradient", "schur_dense", "schur_cg"])
def test_singular_gauss_newton_step_is_rejected(device, solver):
    """Gauss-Newton (no damping) with singular diagonal blocks (J_a^T J_a, and the
    exactly zero block of an unobserved variable): the factorisation -- block-Jacobi
    for CG, the eliminated block inverse for Schur -- reports a non-positive pivot
    (LinAlgError), the step is unusable, and without a trust region the solve stops
    (solvers.py:595-596, 667-693)."""
    print(f"lm singular GN solver={solver}", flush=True)
    P = two_type_problem(device)
    opts = dict(linear_solver=solver, eliminate="b") if solver != "conjugate_gradient" else {}
    _, s = P.solve(return_summary=True, trust_region=None, **opts)
    assert s.termination_reason == "no_acceptable_step"
    assert (s.iterations, s.accepted_steps, s.rejected_steps) == (1, 0, 1)


@pytest.mark.parametrize("solver", ["conjugate_gradient", "schur_dense", "schur_cg"])
def test_damped_solve_with_unobserved_variable(device, solver):
    """With damping the same problem solves: the unobserved blocks are lambda I."""
    P = two_type_problem(device)
    sol, s = P.solve(return_summary=True, linear_solver=solver, **({"eliminate": "b"} if solver != "conjugate_gradient"
                                                                   else {}))
    assert s.final_cost < 1e-12
    # the unobserved variables do not move
    assert np.allclose(sol["a"][2].cpu().numpy(), 1.0) and np.allclose(sol["b"][3].cpu().numpy(), 1.0)
