"""Implicit gradients of the LM optimum vs float64 central finite differences."""

import numpy as np
import pytest
import torch

from torchls_amd import ConjugateGradientConfig, HuberLoss, Problem, SE3Var, TerminationConfig, generators
from torchls_amd.geometry import se3_log
from torchls_amd.implicit import solve_implicit
from torchls_amd.residuals import reprojection

pytestmark = pytest.mark.gpu


def camera_prior(camera, reference):
    return (camera[:3, :] - reference[:3, :]).reshape(-1)


def point_prior(point, reference):
    return point - reference


OPTS = dict(conjugate_gradient=ConjugateGradientConfig(rtol=1e-13, atol=0.0, max_iterations=500),
            termination=TerminationConfig(max_iterations=60, cost_tolerance=0.0, gradient_tolerance=1e-14,
                                          parameter_tolerance=0.0))


def build(device, delta):
    prob = generators.bundle_adjustment(8, 60, 6, seed=31, noise=2e-3)
    obs = torch.tensor(prob.observations, device=device, requires_grad=True)
    P = Problem({"cams": torch.tensor(prob.cameras_init, device=device),
                 "points": torch.tensor(prob.points_init, device=device)}, manifolds={"cams": SE3Var})
    P.add_residual(reprojection, ("cams", torch.tensor(prob.cam_idx)), ("points", torch.tensor(prob.pt_idx)), obs,
                   loss=HuberLoss(delta) if delta else None)
    P.add_residual(camera_prior, ("cams", 0), torch.tensor(prob.cameras_true[0], device=device))
    P.add_residual(point_prior, ("points", 0), torch.tensor(prob.points_true[0], device=device))
    truth = torch.tensor(prob.cameras_true, device=device)
    return prob, P, obs, truth


def loss_fn(cams, truth):
    rel = torch.linalg.solve(truth, cams)
    return se3_log(rel).square().sum()


def polish(P, steps=4):
    """Newton steps with the exact float64 Hessian: solutions accurate to rounding."""
    from torchls_amd.implicit import _objective

    an = P._analyze()
    vals = {k: v.detach().clone() for k, v in P.parameters.items()}
    for _ in range(steps):
        d0 = torch.zeros(an.tangent_dim, dtype=torch.float64, device=next(iter(vals.values())).device)
        f = lambda d: _objective(an, vals, d)  # noqa: E731
        g = torch.func.grad(f)(d0)
        H = torch.func.hessian(f)(d0)
        vals = an.retract(vals, -torch.linalg.solve(H, g))
    return vals


# refine = 0 is the single Gauss-Newton adjoint solve; its error is the Gauss-Newton
# approximation of the Hessian at nonzero residuals (about 2% on this problem).
@pytest.mark.parametrize("delta,refine,tol", [(None, 30, 1e-5), (3e-3, 30, 1e-5), (None, 0, 1e-1)])
def test_adjoint_gradient_matches_finite_differences(device, delta, refine, tol):
    print(f"implicit delta={delta} refine={refine}", flush=True)
    prob, P, obs, truth = build(device, delta)
    sol = solve_implicit(P, obs, refine=refine, **OPTS)
    with torch.no_grad():
        P.parameters.update(polish(P))
    L = loss_fn(sol["cams"], truth)
    (g,) = torch.autograd.grad(L, obs)
    theta = {k: v.detach().clone() for k, v in P.parameters.items()}
    rng = np.random.default_rng(0)
    # probe coordinates away from the Huber kink
    with torch.no_grad():
        res = P.residuals().reshape(-1)[: 2 * prob.n_obs].reshape(-1, 2)
        norms = res.norm(dim=-1).cpu().numpy()
    ok = np.flatnonzero(np.abs(norms - (delta or 1.0)) > 5e-4)
    probes = rng.choice(ok, size=12, replace=False)
    h = 1e-6
    errs = []
    for i in probes:
        c = int(rng.integers(0, 2))
        vals = []
        for sgn in (1.0, -1.0):
            with torch.no_grad():
                obs[i, c] += sgn * h
            P.parameters.update({k: v.clone() for k, v in theta.items()})
            P.solve(**OPTS)
            vals.append(float(loss_fn(polish(P)["cams"], truth)))
            with torch.no_grad():
                obs[i, c] -= sgn * h
        fd = (vals[0] - vals[1]) / (2 * h)
        errs.append((float(g[i, c]), fd))
    a = np.array(errs)
    rel = np.linalg.norm(a[:, 0] - a[:, 1]) / np.linalg.norm(a[:, 1])
    print(f"adjoint vs FD relative error {rel:.3e}", flush=True)
    assert rel <= tol
