"""LM solves through Problem.solve() on the HIP backend vs the float64 oracle LM."""

import math

import numpy as np
import pytest
import torch

from oracle.ba import BAProblem
from oracle.lm import Termination, solve as oracle_solve
from torchls_amd import ConjugateGradientConfig, HuberLoss, Problem, SE3Var, TerminationConfig, generators
from torchls_amd.residuals import reprojection

pytestmark = pytest.mark.gpu


def plain_reprojection(camera, point, observation):
    local = camera[:3, :3] @ point + camera[:3, 3]
    return local[:2] / local[2] - observation


def build(prob, device, fn=reprojection, dtype=torch.float64):
    P = Problem(
        {"cams": torch.tensor(prob.cameras_init, device=device, dtype=dtype),
         "points": torch.tensor(prob.points_init, device=device, dtype=dtype)},
        manifolds={"cams": SE3Var},
    )
    P.add_residual(fn, ("cams", torch.tensor(prob.cam_idx)), ("points", torch.tensor(prob.pt_idx)),
                   torch.tensor(prob.observations, device=device, dtype=dtype), loss=HuberLoss(1e-2))
    return P


def oracle(prob, iters):
    o = BAProblem(prob.cam_idx, prob.pt_idx, prob.observations, prob.n_cameras, prob.n_points, huber_delta=1e-2)
    return oracle_solve(o, (prob.cameras_init, prob.points_init), term=Termination(max_iterations=iters))


@pytest.mark.parametrize("fn", [reprojection, plain_reprojection])
def test_float64_matches_oracle_trajectory(device, fn):
    print(f"lm float64 fn={fn.__name__}", flush=True)
    prob = generators.ba_small()
    P = build(prob, device, fn)
    sol, s = P.solve(return_summary=True, termination=TerminationConfig(max_iterations=20))
    _, ref = oracle(prob, 20)
    assert s.termination_reason == ref.termination_reason
    assert (s.iterations, s.accepted_steps, s.rejected_steps) == (ref.iterations, ref.accepted_steps, ref.rejected_steps)
    assert s.lambda_history == pytest.approx(ref.lambda_history, rel=1e-12)
    assert s.linear_iterations_history == ref.linear_iterations_history
    assert np.allclose(s.cost_history, ref.cost_history, rtol=1e-9, atol=0)
    assert math.isclose(s.final_cost, ref.final_cost, rel_tol=1e-9)
    assert sol["cams"].shape == (16, 4, 4) and P.parameters["points"] is sol["points"]


def test_mixed_precision_final_cost(device):
    prob = generators.ba_small()
    P = build(prob, device)
    cg = ConjugateGradientConfig(matvec_dtype=torch.float32)
    _, s = P.solve(return_summary=True, conjugate_gradient=cg, termination=TerminationConfig(max_iterations=20))
    _, ref = oracle(prob, 20)
    assert abs(s.final_cost - ref.final_cost) <= 1e-6 * ref.final_cost


def test_readme_triangulation(device):
    def reprojection_guarded(point, camera, observed_uv):
        xyz = camera[:3, :3] @ point + camera[:3, 3]
        valid = xyz[2] > 1e-6
        z = torch.where(valid, xyz[2], torch.ones_like(xyz[2]))
        error = xyz[:2] / z - observed_uv
        return torch.where(valid, error, torch.zeros_like(error))

    cameras = torch.eye(4, dtype=torch.float64).repeat(2, 1, 1)
    cameras[1, 0, 3] = -1.0
    obs = torch.tensor([[0.2, 0.1], [0.0, 0.1]], dtype=torch.float64)
    P = Problem({"points": torch.tensor([[0.8, 0.4, 4.0]], dtype=torch.float64, device=device)})
    P.add_residual(reprojection_guarded, ("points", torch.tensor([0, 0])), cameras.to(device), obs.to(device))
    sol = P.solve()
    assert torch.allclose(sol["points"].cpu(), torch.tensor([[1.0, 0.5, 5.0]], dtype=torch.float64), atol=1e-8)


GOLDEN_CASES = {
    "cg": dict(),
    "cg_adaptive": dict(trust_region="adaptive"),
    "schur_cg": dict(linear_solver="schur_cg", eliminate="points"),
    "schur_dense": dict(linear_solver="schur_dense", eliminate="points"),
}


@pytest.mark.parametrize("name", list(GOLDEN_CASES))
def test_hip_lm_against_reference_golden(device, name):
    """HIP LM (warp_pcg control flow) vs the reference's own torch-backend trajectory."""
    import os

    from torchls_amd import TrustRegionConfig

    print(f"golden lm {name}", flush=True)
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "lm.npz"), allow_pickle=False)
    P = Problem({"cams": torch.tensor(g["cams0"], device=device), "points": torch.tensor(g["pts0"], device=device)},
                manifolds={"cams": SE3Var})
    P.add_residual(reprojection, ("cams", torch.tensor(g["cam_idx"])), ("points", torch.tensor(g["pt_idx"])),
                   torch.tensor(g["uv"], device=device), loss=HuberLoss(3e-3))
    opts = dict(GOLDEN_CASES[name])
    if opts.pop("trust_region", None) == "adaptive":
        opts["trust_region"] = TrustRegionConfig(strategy="adaptive")
    sol, s = P.solve(return_summary=True, termination=TerminationConfig(max_iterations=20), **opts)
    assert s.termination_reason == str(g[name + "_reason"])
    assert abs(s.iterations - int(g[name + "_counts"][0])) <= 1
    # while the predicted reductions are well above rounding (the first five
    # iterations here; the cost then sits on a plateau where acceptance is decided
    # by the last bits) the damping and PCG counts are the reference's own
    assert s.lambda_history[:6] == pytest.approx(list(g[name + "_lambda"][:6]), rel=1e-12)
    assert s.linear_iterations_history[:5] == list(g[name + "_linear"][:5])
    assert np.allclose(s.cost_history[:6], g[name + "_cost"][:6], rtol=1e-10, atol=0)
    assert math.isclose(s.final_cost, float(g[name + "_cost"][-1]), rel_tol=1e-8)
    assert np.allclose(sol["points"].cpu().numpy(), g[name + "_pts"], rtol=0, atol=1e-6)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_step_finiteness_flag_propagates_nan_and_inf(device, dtype):
    """The LM loop reads max |step| (one float64 reduction) as the step's finiteness flag."""
    import math

    print(f"finiteness flag dtype={dtype}", flush=True)
    x = torch.randn(1_000_003, dtype=dtype, device=device)
    assert math.isfinite(torch.linalg.vector_norm(x, ord=math.inf, dtype=torch.float64).item())
    for pos in (0, 517_231, 1_000_002):
        for bad in (float("nan"), float("inf"), -float("inf")):
            y = x.clone()
            y[pos] = bad
            assert not math.isfinite(torch.linalg.vector_norm(y, ord=math.inf, dtype=torch.float64).item())


@pytest.mark.parametrize("solver", ["conjugate_gradient", "schur_cg"])
@pytest.mark.parametrize("mode", ["jacobi", "diagonal", "both"])
def test_column_scaling_matches_oracle(device, solver, mode):
    """jacobi_scaling and damping_mode="diagonal" (solvers.py:430-472): the scale is
    applied while packing the blocks; the LM trajectory equals the oracle's."""
    from oracle.lm import TrustRegion
    from torchls_amd import TrustRegionConfig

    print(f"lm scaling solver={solver} mode={mode}", flush=True)
    prob = generators.ba_small()
    P = build(prob, device)
    damping_mode = "diagonal" if mode in ("diagonal", "both") else "identity"
    opts = dict(linear_solver=solver, eliminate="points") if solver == "schur_cg" else {}
    _, s = P.solve(return_summary=True, termination=TerminationConfig(max_iterations=8),
                   trust_region=TrustRegionConfig(damping_mode=damping_mode),
                   jacobi_scaling=mode in ("jacobi", "both"), **opts)
    o = BAProblem(prob.cam_idx, prob.pt_idx, prob.observations, prob.n_cameras, prob.n_points, huber_delta=1e-2)
    _, ref = oracle_solve(o, (prob.cameras_init, prob.points_init), term=Termination(max_iterations=8),
                          tr=TrustRegion(damping_mode=damping_mode), jacobi_scaling=mode in ("jacobi", "both"),
                          linear_solver=solver, eliminate=1 if solver == "schur_cg" else None)
    assert s.termination_reason == ref.termination_reason
    assert (s.iterations, s.accepted_steps, s.rejected_steps) == (ref.iterations, ref.accepted_steps,
                                                                  ref.rejected_steps)
    assert s.linear_iterations_history == ref.linear_iterations_history
    assert s.lambda_history == pytest.approx(ref.lambda_history, rel=1e-12)
    assert math.isclose(s.final_cost, ref.final_cost, rel_tol=1e-9)


def test_lm_decide_matches_oracle(device):
    """tlh_lm_decide (gain ratio, acceptance, damping update on the device) equals the
    oracle's trial_decision on hand-picked edge cases and 300 random trials per
    trust-region rule, with and without a usable step: bit for bit, except that an
    accepted adaptive step's damping may differ by up to 4 ulps (the multiplier's cube
    is rounded once on the device, the host's libm pow to about half an ulp; one ulp of
    a cube of magnitude <= 1 moves 1 - cube by <= 2.2e-16, <= 3 ulps of a multiplier
    >= 1/3, and the product rounds once more)."""
    print("lm decide vs oracle", flush=True)
    import random

    from oracle.lm import TrustRegion, trial_decision
    from torchls_amd import _lib
    from torchls_amd.hip_ops import lm_decide
    rules = [TrustRegion(lambda_min=1e-9, lambda_max=1e6, lambda_factor=4.0, step_quality_min=1e-4),
             TrustRegion(lambda_min=1e-9, lambda_max=1e6, step_quality_min=1e-4, strategy="adaptive"), None]
    rng = random.Random(11)
    cases = [(1.0, 0.9, 1.0), (1.0, 1.0, 1.0), (1.0, -1.0, -2.0), (1.0, 0.0, 1.0), (math.nan, 1.0, 1.0),
             (math.inf, 1.0, 1.0), (1.0, 5.0, 1.0), (1.0, 1e-5, 1.0)]
    cases += [(rng.uniform(0, 2), rng.uniform(-0.5, 2.0), rng.uniform(-0.2, 2.0)) for _ in range(300)]
    for tr in rules:
        trust = _lib.LmTrust(1 if tr else 0, 1 if tr and tr.strategy == "adaptive" else 0,
                             tr.lambda_min if tr else 0.0, tr.lambda_max if tr else 0.0,
                             tr.lambda_factor if tr else 0.0, tr.step_quality_min if tr else 0.0)
        for k, (obj, red, pred) in enumerate(cases):
            lam = (0.0 if tr is None else [1e-3, 1e6, 3.7e-2][k % 3])
            rej = [2.0, 8.0][k % 2]
            usable = k % 11 != 5
            state = torch.tensor([lam, rej], dtype=torch.float64, device=device)
            if usable:
                t = [torch.tensor([v], dtype=torch.float64, device=device) for v in (obj, red, 0.0)]
                stats = torch.tensor([-pred, 4.0, 0.0, 0.0], dtype=torch.float64, device=device)  # -g.d = pred
                out = lm_decide((t[0], t[1], t[2], stats, torch.zeros(1, dtype=torch.float64, device=device)),
                                        state, trust).cpu().tolist()
            else:
                out = lm_decide(None, state, trust).cpu().tolist()
            q, acc, lam_n, rej_n, stop = trial_decision(obj, red, pred, lam, rej, tr, usable=usable)
            st = state.cpu().tolist()
            assert (out[2] == 1.0) == acc and (out[3] == 1.0) == usable and (out[6] == 1.0) == stop
            tol = 4 * math.ulp(lam_n) if (acc and tr is not None and tr.strategy == "adaptive") else 0.0
            assert abs(out[7] - lam_n) <= tol and abs(st[0] - lam_n) <= tol and out[7] == st[0]
            assert out[8] == rej_n and st[1] == rej_n and out[5] == lam
            if usable:
                assert out[0] == pred and out[4] == 2.0
                assert out[1] == q or (math.isnan(out[1]) and math.isnan(q))
