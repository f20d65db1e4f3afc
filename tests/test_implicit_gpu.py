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


@pytest.mark.parametrize("delta,weighted", [(None, False), (3e-3, True)])
def test_exact_hessian_kernel_matches_torch_func(device, delta, weighted):
    """tlh_reprojection_hvp + tlh_segment_sum_f64 (hyper-dual numbers) against torch.func's
    forward-over-reverse Hessian-vector product of the same float64 objective through the
    SE(3) retraction (Rodrigues exponential), away from the optimum and at it."""
    from torchls_amd.implicit import ExactHessian, _objective

    print(f"exact hvp delta={delta} weighted={weighted}", flush=True)
    prob = generators.bundle_adjustment(8, 60, 6, seed=31, noise=2e-3)
    obs = torch.tensor(prob.observations, device=device)
    P = Problem({"cams": torch.tensor(prob.cameras_init, device=device),
                 "points": torch.tensor(prob.points_init, device=device)}, manifolds={"cams": SE3Var})
    w = torch.rand(prob.n_obs, dtype=torch.float64, generator=torch.Generator().manual_seed(2)) + 0.5
    P.add_residual(reprojection, ("cams", torch.tensor(prob.cam_idx)), ("points", torch.tensor(prob.pt_idx)), obs,
                   loss=HuberLoss(delta) if delta else None, weights=w if weighted else None)
    an = P._analyze()
    rng = np.random.default_rng(1)
    for vals in (P.parameters, P.solve(**OPTS)):
        H = ExactHessian(an, vals, damping=0.25)
        assert len(H.hip) == 1 and not H.others  # the reprojection batch runs in HIP
        for _ in range(3):
            v = torch.tensor(rng.standard_normal(an.tangent_dim), device=device)
            zero = torch.zeros_like(v)
            ref = torch.func.jvp(torch.func.grad(lambda d: _objective(an, vals, d)), (zero,), (v,))[1] + 0.25 * v
            got = H(v)
            err = float((got - ref).norm() / ref.norm())
            assert err <= 1e-11, err


def test_gradient_matches_dense_qr_unrolled_solve(device):
    """The implicit gradient against torch.func-style reverse differentiation through an
    unrolled Gauss-Newton solve whose steps are dense QR least-squares solves (the dense
    path of solvers.py:504-510: [J; sqrt(lambda) I] delta = [-r; 0], here lambda = 0 with
    the gauge fixed by priors), written here in plain torch."""
    from torchls_amd.geometry import se3_exp

    print("implicit vs dense-QR unrolled", flush=True)
    prob, P, obs, truth = build(device, None)
    sol = solve_implicit(P, obs, refine=30, polish=2, **OPTS)
    (g_impl,) = torch.autograd.grad(loss_fn(sol["cams"], truth), obs)

    ci, pi = torch.tensor(prob.cam_idx, device=device), torch.tensor(prob.pt_idx, device=device)
    ref_c = torch.tensor(prob.cameras_true[0], device=device)
    ref_p = torch.tensor(prob.points_true[0], device=device)
    nc, npt = prob.n_cameras, prob.n_points

    def residual(cams, pts, uv):
        local = (cams[ci, :3, :3] @ pts[pi][..., None])[..., 0] + cams[ci, :3, 3]
        r = local[:, :2] / local[:, 2:3] - uv
        return torch.cat((r.reshape(-1), (cams[0, :3, :] - ref_c[:3, :]).reshape(-1), pts[0] - ref_p))

    def moved(cams, pts, d):
        return cams @ se3_exp(d[: 6 * nc].reshape(nc, 6)), pts + d[6 * nc:].reshape(npt, 3)

    x_obs = obs.detach().clone().requires_grad_(True)
    cams = torch.tensor(prob.cameras_init, device=device)
    pts = torch.tensor(prob.points_init, device=device)
    zero = torch.zeros(6 * nc + 3 * npt, dtype=torch.float64, device=device)
    for _ in range(25):
        J = torch.func.jacrev(lambda d: residual(*moved(cams, pts, d), x_obs))(zero)
        r = residual(cams, pts, x_obs)
        Q, R = torch.linalg.qr(J)
        d = torch.linalg.solve_triangular(R, -(Q.T @ r)[:, None], upper=True)[:, 0]
        cams, pts = moved(cams, pts, d)
    (g_qr,) = torch.autograd.grad(loss_fn(cams, truth), x_obs)
    err = float((g_impl - g_qr).norm() / g_qr.norm())
    print(f"  implicit vs unrolled dense QR: rel err {err:.2e}", flush=True)
    assert err <= 1e-6


def test_adjoint_gradient_fd_differentiable_config(device):
    """500 cameras, 100,000 points, 1,000,000 observations (float32 matvec forward solve,
    then Newton polishing on the exact float64 Hessian): the implicit gradient against
    float64 central differences (h = 1e-6) on 64 random observation coordinates, each
    perturbed optimum polished by exact-Hessian Newton steps."""
    from torchls_amd.implicit import newton_polish

    prob = generators.ba_differentiable()
    print(f"implicit FD at {prob.n_obs} observations", flush=True)
    obs = torch.tensor(prob.observations, device=device, requires_grad=True)
    P = Problem({"cams": torch.tensor(prob.cameras_init, device=device),
                 "points": torch.tensor(prob.points_init, device=device)}, manifolds={"cams": SE3Var})
    P.add_residual(reprojection, ("cams", torch.tensor(prob.cam_idx)), ("points", torch.tensor(prob.pt_idx)), obs,
                   loss=HuberLoss(1e-2))
    P.add_residual(camera_prior, ("cams", 0), torch.tensor(prob.cameras_true[0], device=device))
    P.add_residual(point_prior, ("points", 0), torch.tensor(prob.points_true[0], device=device))
    truth = torch.tensor(prob.cameras_true, device=device)
    opts = dict(conjugate_gradient=ConjugateGradientConfig(matvec_dtype=torch.float32, stop_on_stagnation=True),
                termination=TerminationConfig(max_iterations=10))
    sol = solve_implicit(P, obs, refine=30, polish=2, **opts)
    print("  forward solve and polish done", flush=True)
    L = loss_fn(sol["cams"], truth)
    (g,) = torch.autograd.grad(L, obs)
    print("  adjoint done", flush=True)
    theta = {k: v.detach().clone() for k, v in P.parameters.items()}
    with torch.no_grad():
        norms = P.residuals().reshape(-1)[: 2 * prob.n_obs].reshape(-1, 2).norm(dim=-1).cpu().numpy()
    rng = np.random.default_rng(0)
    ok = np.flatnonzero(np.abs(norms - 1e-2) > 5e-4)  # away from the Huber kink
    probes = rng.choice(ok, size=64, replace=False)
    h = 1e-6
    pairs = []
    import time

    t0 = time.perf_counter()
    for n, i in enumerate(probes):
        c = int(rng.integers(0, 2))
        vals = []
        for sgn in (1.0, -1.0):
            with torch.no_grad():
                obs[i, c] += sgn * h
            P.parameters.update({k: v.clone() for k, v in theta.items()})
            # from the unperturbed optimum the perturbed one is O(h) away: two Newton steps
            # (quadratic convergence) reach it to rounding with solves to rtol 1e-10
            vals.append(float(loss_fn(newton_polish(P, steps=2, rtol=1e-10)["cams"], truth)))
            with torch.no_grad():
                obs[i, c] -= sgn * h
        pairs.append((float(g[i, c]), (vals[0] - vals[1]) / (2 * h)))
        print(f"  probe {n}: obs[{i}, {c}] adjoint {pairs[-1][0]:.9e} FD {pairs[-1][1]:.9e} "
              f"({time.perf_counter() - t0:.1f} s)", flush=True)
    a = np.array(pairs)
    rel = np.linalg.norm(a[:, 0] - a[:, 1]) / np.linalg.norm(a[:, 1])
    print(f"  64 probes: adjoint vs FD relative error {rel:.3e}", flush=True)
    assert rel <= 1e-5
