"""The general Schur path (structures outside the fused bundle-adjustment kernels) against
the oracle's Schur closed form, and its LM solves against the full-system PCG solve."""

import math

import numpy as np
import pytest
import torch

from oracle.schur import SchurSystem
from oracle.sparse import NpGroup, NpJacobian
from tests.helpers import rel_err
from torchls_amd import HuberLoss, Problem, SE3Var, TerminationConfig, generators
from torchls_amd.residuals import reprojection
from torchls_amd.schur import HipSchurPlan

pytestmark = pytest.mark.gpu


def project(camera, point):
    local = camera[:3, :3] @ point + camera[:3, 3]
    return local[:2] / local[2]


def stereo(left, right, point, uv):
    """One point seen by two cameras in one residual: two retained slots of one type."""
    return torch.cat((project(left, point), project(right, point))) - uv


def focal_reprojection(camera, focal, point, uv):
    """A per-camera focal length as its own retained type: retained-retained coupling."""
    return focal[0] * project(camera, point) - uv


def stereo_problem(device, seed=3):
    prob = generators.bundle_adjustment(12, 150, 4, seed=seed)
    rng = np.random.default_rng(seed)
    left = rng.integers(0, prob.n_cameras, prob.n_points)
    right = (left + 1 + rng.integers(0, prob.n_cameras - 1, prob.n_points)) % prob.n_cameras
    pts = torch.tensor(prob.points_true)
    cams = torch.tensor(prob.cameras_true)
    uv = torch.stack([torch.cat((project(cams[a], p), project(cams[b], p)))
                      for a, b, p in zip(left, right, pts)]) + 1e-3 * torch.randn(prob.n_points, 4, dtype=torch.float64,
                                                                                   generator=torch.Generator().manual_seed(1))
    P = Problem({"cams": torch.tensor(prob.cameras_init, device=device),
                 "points": torch.tensor(prob.points_init, device=device)}, manifolds={"cams": SE3Var})
    P.add_residual(stereo, ("cams", torch.tensor(left)), ("cams", torch.tensor(right)),
                   ("points", torch.arange(prob.n_points)), uv.to(device), loss=HuberLoss(1e-2))
    # gauge: a prior on the first camera
    P.add_residual(lambda c, r: (c[:3, :] - r[:3, :]).reshape(-1), ("cams", 0),
                   torch.tensor(prob.cameras_true[0], device=device))
    return P


def focal_problem(device, seed=4):
    prob = generators.bundle_adjustment(10, 200, 5, seed=seed)
    P = Problem({"cams": torch.tensor(prob.cameras_init, device=device),
                 "focal": torch.full((prob.n_cameras, 1), 1.02, dtype=torch.float64, device=device),
                 "points": torch.tensor(prob.points_init, device=device)}, manifolds={"cams": SE3Var})
    ci = torch.tensor(prob.cam_idx)
    P.add_residual(focal_reprojection, ("cams", ci), ("focal", ci), ("points", torch.tensor(prob.pt_idx)),
                   torch.tensor(prob.observations, device=device))
    P.add_residual(lambda f: f - 1.0, ("focal", torch.arange(prob.n_cameras)))
    P.add_residual(lambda c, r: (c[:3, :] - r[:3, :]).reshape(-1), ("cams", 0),
                   torch.tensor(prob.cameras_true[0], device=device))
    return P


def split_problem(device):
    """Bundle adjustment whose observations come in two residual groups (two coupling groups)."""
    prob = generators.bundle_adjustment(16, 300, 6, seed=5)
    half = prob.n_obs // 2
    P = Problem({"cams": torch.tensor(prob.cameras_init, device=device),
                 "points": torch.tensor(prob.points_init, device=device)}, manifolds={"cams": SE3Var})
    for sl in (slice(0, half), slice(half, None)):
        P.add_residual(reprojection, ("cams", torch.tensor(prob.cam_idx[sl])), ("points", torch.tensor(prob.pt_idx[sl])),
                       torch.tensor(prob.observations[sl], device=device), loss=HuberLoss(1e-2))
    return P


PROBLEMS = {"stereo": stereo_problem, "focal": focal_problem, "split": split_problem}


def to_np(jacobian, layout):
    groups = [NpGroup([b.double().cpu().numpy() for b in g.blocks], list(g.type_indices),
                      [i.cpu().numpy() for i in g.variable_indices]) for g in jacobian.groups]
    return NpJacobian(groups, [(e.count, e.dim, e.offset) for e in layout])


@pytest.mark.parametrize("name", list(PROBLEMS))
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-11), (torch.float32, 2e-5)])
def test_general_schur_operators(device, name, dtype, tol):
    print(f"general schur operators {name} dtype={dtype}", flush=True)
    P = PROBLEMS[name](device)
    an = P._analyze()
    lin = an.linearize(P.parameters)
    diag = tuple(d.double() for d in lin.jacobian.diagonal_blocks())
    lam = 1e-2
    plan = HipSchurPlan(an, "points", device, dtype=dtype)
    assert plan.general
    system = plan.linearize(lin.jacobian, diag, dense=True).damped(lam)
    J = to_np(lin.jacobian, an.layout)
    e = [k for k, ent in enumerate(an.layout) if ent.name == "points"][0]
    ref = SchurSystem(J, [d.cpu().numpy() for d in diag], e, lam)
    rng = np.random.default_rng(3)
    rhs = rng.standard_normal(J.n)
    assert rel_err(system.reduced_rhs(torch.tensor(rhs, device=device)).cpu().numpy(), ref.reduced_rhs(rhs)) <= tol
    x = torch.tensor(rng.standard_normal(ref.dim), device=device, dtype=dtype)
    xd = x.double().cpu().numpy()
    assert rel_err(system.matvec(x).cpu().numpy(), ref.matvec(xd)) <= tol
    assert rel_err(system.backsubstitute(x).cpu().numpy(), ref.backsubstitute(xd)) <= tol
    for a, b in zip(system.preconditioner_blocks(), ref.preconditioner_blocks()):
        assert rel_err(a.cpu().numpy(), b) <= tol
    assert rel_err(np.tril(system.dense_matrix().cpu().numpy()), np.tril(ref.dense_matrix())) <= tol


@pytest.mark.parametrize("name", list(PROBLEMS))
@pytest.mark.parametrize("solver", ["schur_cg", "schur_dense"])
def test_general_schur_lm_reaches_the_full_system_optimum(device, name, solver):
    print(f"general schur lm {name} {solver}", flush=True)
    term = TerminationConfig(max_iterations=15)
    P0 = PROBLEMS[name](device)
    _, ref = P0.solve(return_summary=True, termination=term)
    P = PROBLEMS[name](device)
    _, s = P.solve(return_summary=True, linear_solver=solver, eliminate="points", termination=term)
    print(f"  full-system {ref.final_cost:.12e} ({ref.termination_reason}), {solver} {s.final_cost:.12e} "
          f"({s.termination_reason})", flush=True)
    assert s.final_cost < s.initial_cost
    assert math.isclose(s.final_cost, ref.final_cost, rel_tol=1e-8)
