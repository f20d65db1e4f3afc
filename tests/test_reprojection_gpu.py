"""HIP reprojection linearization and fused trial objective vs the float64 oracle."""

import numpy as np
import pytest
import torch

from oracle.ba import huber_reduction, huber_value, linearize_ba, reprojection as oracle_reprojection
from tests.helpers import rel_err
from torchls_amd import HuberLoss, Problem, SE3Var, generators
from torchls_amd.residuals import reprojection

pytestmark = pytest.mark.gpu


def problem(prob, device, dtype, loss, weights, rows=4):
    cams = torch.tensor(prob.cameras_init, dtype=dtype, device=device)[:, :rows].contiguous()
    P = Problem({"cams": cams, "points": torch.tensor(prob.points_init, dtype=dtype, device=device)},
                manifolds={"cams": SE3Var})
    P.add_residual(reprojection, ("cams", torch.tensor(prob.cam_idx)), ("points", torch.tensor(prob.pt_idx)),
                   torch.tensor(prob.observations, dtype=dtype, device=device), loss=loss, weights=weights)
    return P


@pytest.mark.parametrize("rows", [4, 3])
@pytest.mark.parametrize("delta", [None, 1e-2])
@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("dtype,tol,rtol", [(torch.float64, 1e-12, 1e-12), (torch.float32, 1e-5, 1e-4)])
def test_linearize_matches_oracle(device, rows, delta, weighted, dtype, tol, rtol):
    """Blocks within ``tol``; residual-derived values within ``rtol`` (a float32
    residual is a difference of O(1) projections, so it carries ~1e-5 relative
    rounding at the generator's noise level).  The oracle reads the same
    (rounded) inputs."""
    print(f"reprojection linearize rows={rows} delta={delta} weighted={weighted} dtype={dtype}", flush=True)
    prob = generators.ba_small()
    rng = np.random.default_rng(3)
    w = rng.uniform(0.2, 2.0, len(prob.cam_idx)) if weighted else None
    P = problem(prob, device, dtype, None if delta is None else HuberLoss(delta), w, rows)
    an = P._analyze()
    assert an.batches[0].hip_operands(P.parameters, an.layout, device, dtype) is not None
    lin = an.linearize(P.parameters)
    cams, pts, obs = (P.parameters["cams"].double().cpu().numpy(), P.parameters["points"].double().cpu().numpy(),
                      torch.tensor(prob.observations, dtype=dtype).double().numpy())
    if rows == 3:
        cams = np.concatenate((cams, np.tile(np.array([[[0.0, 0.0, 0.0, 1.0]]]), (len(cams), 1, 1))), axis=1)
    model, Jc, Jp, raw, half_obj = linearize_ba(cams, pts, prob.cam_idx, prob.pt_idx, obs, delta)
    sw = np.ones(len(raw)) if w is None else np.sqrt(w)
    g = lin.jacobian.groups[0]
    assert rel_err(g.blocks[0].cpu().numpy(), Jc * sw[:, None, None]) <= tol
    assert rel_err(g.blocks[1].cpu().numpy(), Jp * sw[:, None, None]) <= tol
    assert rel_err(lin.residual.cpu().numpy().reshape(-1, 2), model * sw[:, None]) <= rtol
    if delta is not None or weighted:
        assert rel_err(lin.raw_residual.cpu().numpy().reshape(-1, 2), raw) <= rtol
        s = np.sum(raw * raw, -1)
        value = s if delta is None else huber_value(s, delta)
        ref = 0.5 * float(np.sum(value * (1.0 if w is None else w)))
        assert abs(float(lin.objective) - ref) <= rtol * ref
        if w is None:
            assert abs(0.5 * float(np.sum(value)) - half_obj) <= 1e-12 * half_obj
    else:
        assert lin.raw_residual is None and lin.objective is None


@pytest.mark.parametrize("delta", [None, 1e-2])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 1e-4)])
def test_trial_matches_oracle(device, delta, dtype, tol):
    print(f"reprojection trial delta={delta} dtype={dtype}", flush=True)
    prob = generators.ba_small()
    P = problem(prob, device, dtype, None if delta is None else HuberLoss(delta), None)
    an = P._analyze()
    lin = an.linearize(P.parameters)
    rng = np.random.default_rng(5)
    step = torch.tensor(1e-3 * rng.standard_normal(an.tangent_dim), dtype=torch.float64, device=device)
    moved = an.retract(P.parameters, step)
    objective, reduction = an.evaluate_trial(moved, lin)
    cams = moved["cams"].double().cpu().numpy()
    pts = moved["points"].double().cpu().numpy()
    obs = torch.tensor(prob.observations, dtype=dtype).double().numpy()
    before, _, _ = oracle_reprojection(P.parameters["cams"].double().cpu().numpy(),
                                       P.parameters["points"].double().cpu().numpy(), prob.cam_idx, prob.pt_idx, obs)
    after, _, _ = oracle_reprojection(cams, pts, prob.cam_idx, prob.pt_idx, obs)
    s0, s1 = np.sum(before ** 2, -1), np.sum(after ** 2, -1)
    diff = np.sum((before - after) * (before + after), -1)
    if delta is None:
        ref_obj, ref_red = 0.5 * s1.sum(), 0.5 * diff.sum()
    else:
        ref_obj = 0.5 * huber_value(s1, delta).sum()
        ref_red = 0.5 * huber_reduction(s0, s1, diff, delta).sum()
    assert abs(float(objective) - ref_obj) <= tol * ref_obj
    # the reduction is a small difference of large sums: compare against the objective's scale
    assert abs(float(reduction) - ref_red) <= tol * ref_obj
    assert abs(float(an.cost(moved)) - ref_obj) <= tol * ref_obj


def test_unsupported_forms_use_the_generic_path(device):
    prob = generators.ba_small()

    def user_fn(camera, point, observation):
        return reprojection(camera, point, observation)

    P = Problem({"cams": torch.tensor(prob.cameras_init, device=device),
                 "points": torch.tensor(prob.points_init, device=device)}, manifolds={"cams": SE3Var})
    P.add_residual(user_fn, ("cams", torch.tensor(prob.cam_idx)), ("points", torch.tensor(prob.pt_idx)),
                   torch.tensor(prob.observations, device=device), loss=HuberLoss(1e-2))
    an = P._analyze()
    assert an.batches[0].hip_operands(P.parameters, an.layout, device, torch.float64) is None
    Q = problem(prob, device, torch.float64, HuberLoss(1e-2), None)
    a = an.linearize(P.parameters)
    b = Q._analyze().linearize(Q.parameters)
    for x, y in zip(a.jacobian.groups[0].blocks, b.jacobian.groups[0].blocks):
        assert rel_err(x.cpu().numpy(), y.cpu().numpy()) <= 1e-12
    assert abs(float(a.objective) - float(b.objective)) <= 1e-12 * float(b.objective)
