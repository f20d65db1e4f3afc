"""Parity at the benchmark sizes: the 5,012,836-observation normal and Schur
operators, the 500,000-observation Schur operators and Schur-PCG LM, the
256-camera dense reduced matrix, and the 10,000-pose graph -- each against the
float64 numpy oracle (or, where the oracle's own closed form is too slow at this
size, a vectorised restatement written here)."""

import math

import numpy as np
import pytest
import torch

from oracle.ba import BAProblem
from oracle.lm import Termination, solve as oracle_solve
from oracle.schur import SchurSystem
from oracle.sparse import NpGroup, NpJacobian
from tests.helpers import rel_err
from torchls_amd import ConjugateGradientConfig, HuberLoss, Problem, SE3Var, TerminationConfig, generators
from torchls_amd.hip_normal import HipNormalPlan
from torchls_amd.residuals import reprojection
from torchls_amd.schur import HipSchurPlan
from torchls_amd.workloads import ba_linearization, pose_graph_problem

pytestmark = pytest.mark.gpu

_CACHE = {}


def workload(name):
    if name not in _CACHE:
        _CACHE.clear()  # one large workload in host memory at a time
        _CACHE[name] = getattr(generators, name)()
    return _CACHE[name]


def ba_problem(prob, device):
    P = Problem({"cams": torch.tensor(prob.cameras_init, device=device),
                 "points": torch.tensor(prob.points_init, device=device)}, manifolds={"cams": SE3Var})
    P.add_residual(reprojection, ("cams", torch.tensor(prob.cam_idx)), ("points", torch.tensor(prob.pt_idx)),
                   torch.tensor(prob.observations, device=device), loss=HuberLoss(1e-2))
    return P


def to_np(jacobian, layout, round_to=None):
    def arr(b):
        b = b if round_to is None else b.to(round_to)
        return b.double().cpu().numpy()

    groups = [NpGroup([arr(b) for b in g.blocks], list(g.type_indices),
                      [i.cpu().numpy() for i in g.variable_indices]) for g in jacobian.groups]
    return NpJacobian(groups, [(e.count, e.dim, e.offset) for e in layout])


@pytest.mark.parametrize("dtype,tol", [(torch.float32, 1e-5), (torch.float64, 1e-12)])
def test_normal_matvec_large(device, dtype, tol):
    """c3: 2,000 cameras, 500,000 points, 5,012,836 observations (the wave kernel)."""
    prob = workload("ba_large")
    print(f"normal matvec 5M dtype={dtype} obs={prob.n_obs}", flush=True)
    _, BJ = ba_linearization(prob, device)
    op = HipNormalPlan(BJ, dtype=dtype).bind(BJ)
    x = np.random.default_rng(0).standard_normal(BJ.shape[1])
    xt = torch.tensor(x, device=device, dtype=dtype)
    y = op.normal_matvec(xt, 1e-3).double().cpu().numpy()
    # the float32 path multiplies float32 blocks and vector: the oracle gets the same rounded inputs
    J = to_np(BJ, BJ.layout, round_to=dtype)
    ref = J.normal_matvec(xt.double().cpu().numpy(), 1e-3)
    err = rel_err(y, ref)
    print(f"  rel err {err:.3e}", flush=True)
    assert err <= tol
    # bitwise repeatable (fixed reduction order)
    assert torch.equal(op.normal_matvec(xt, 1e-3), op.normal_matvec(xt, 1e-3))


@pytest.mark.parametrize("name", ["ba_medium", "ba_large"])
@pytest.mark.parametrize("dtype,tol", [(torch.float32, 2e-5), (torch.float64, 1e-11)])
def test_schur_operators_large(device, name, dtype, tol):
    """Implicit Schur matvec, reduced right-hand side, back-substitution and the
    camera preconditioner blocks at 500,000 and 5,012,836 observations."""
    prob = workload(name)
    print(f"schur operators {name} dtype={dtype} obs={prob.n_obs}", flush=True)
    P = ba_problem(prob, device)
    an = P._analyze()
    lin = an.linearize(P.parameters)
    diag = tuple(d.double() for d in lin.jacobian.diagonal_blocks())
    lam = 1e-3
    system = HipSchurPlan(an, "points", device, dtype=dtype).linearize(lin.jacobian, diag).damped(lam)
    J = to_np(lin.jacobian, an.layout, round_to=dtype)
    ref = SchurSystem(J, [d.cpu().numpy() for d in diag], 1, lam)
    rng = np.random.default_rng(3)
    rhs = rng.standard_normal(J.n)
    rr = system.reduced_rhs(torch.tensor(rhs, device=device))
    e = rel_err(rr.cpu().numpy(), ref.reduced_rhs(rhs))
    x = torch.tensor(rng.standard_normal(ref.dim), device=device, dtype=dtype)
    xd = x.double().cpu().numpy()
    ey = rel_err(system.matvec(x).cpu().numpy(), ref.matvec(xd))
    eb = rel_err(system.backsubstitute(x).cpu().numpy(), ref.backsubstitute(xd))
    ep = max(rel_err(a.cpu().numpy(), b) for a, b in zip(system.preconditioner_blocks(), ref.preconditioner_blocks()))
    print(f"  rel err rhs {e:.2e} matvec {ey:.2e} backsub {eb:.2e} precond {ep:.2e}", flush=True)
    assert max(e, ey, eb, ep) <= tol


def test_schur_pcg_lm_medium(device):
    """c2: Schur-PCG LM at 500,000 observations, float64, against the oracle LM
    (same warp_pcg control flow): iterations, PCG counts, damping and costs; the
    mixed-precision solve reaches the same final cost within 1e-6."""
    prob = workload("ba_medium")
    steps = 3
    print(f"schur-pcg lm 500k obs, {steps} iterations", flush=True)
    term = TerminationConfig(max_iterations=steps)
    _, s = ba_problem(prob, device).solve(return_summary=True, linear_solver="schur_cg", eliminate="points",
                                          termination=term)
    o = BAProblem(prob.cam_idx, prob.pt_idx, prob.observations, prob.n_cameras, prob.n_points, huber_delta=1e-2)
    _, ref = oracle_solve(o, (prob.cameras_init, prob.points_init), linear_solver="schur_cg", eliminate=1,
                          term=Termination(max_iterations=steps))
    print(f"  hip {s.linear_iterations_history} {s.final_cost:.12e}; oracle {ref.linear_iterations_history} "
          f"{ref.final_cost:.12e}", flush=True)
    assert (s.iterations, s.accepted_steps, s.rejected_steps) == (ref.iterations, ref.accepted_steps,
                                                                  ref.rejected_steps)
    assert s.linear_iterations_history == ref.linear_iterations_history
    assert s.lambda_history == pytest.approx(ref.lambda_history, rel=1e-12)
    assert np.allclose(s.cost_history, ref.cost_history, rtol=1e-9, atol=0)
    cg = ConjugateGradientConfig(matvec_dtype=torch.float32, stop_on_stagnation=True)
    _, m = ba_problem(prob, device).solve(return_summary=True, linear_solver="schur_cg", eliminate="points",
                                          termination=term, conjugate_gradient=cg)
    print(f"  mixed {m.linear_iterations_history} {m.final_cost:.12e}", flush=True)
    assert abs(m.final_cost - ref.final_cost) <= 1e-6 * ref.final_cost


def test_schur_pcg_lm_large(device):
    """c3: Schur-PCG LM at 5,012,836 observations, three iterations, float64 and mixed
    precision from the same start: the same iteration and acceptance counts, final
    costs within 1e-6 of each other, and each final cost equal (1e-9) to the objective
    the oracle recomputes on the CPU at the returned cameras and points."""
    prob = workload("ba_large")
    steps = 3
    print(f"schur-pcg lm 5M obs, {steps} iterations", flush=True)
    term = TerminationConfig(max_iterations=steps)
    o = BAProblem(prob.cam_idx, prob.pt_idx, prob.observations, prob.n_cameras, prob.n_points, huber_delta=1e-2)
    out = {}
    for mv in (None, torch.float32):
        opts = dict(stop_on_stagnation=True, max_refinements=0) if mv is not None else {}
        vals, s = ba_problem(prob, device).solve(
            return_summary=True, linear_solver="schur_cg", eliminate="points", termination=term,
            conjugate_gradient=ConjugateGradientConfig(matvec_dtype=mv, **opts))
        cost = o.cost((vals["cams"].cpu().numpy(), vals["points"].cpu().numpy()))
        print(f"  matvec {mv}: {s.termination_reason}, pcg {s.linear_iterations_history}, "
              f"final {s.final_cost:.12e}, oracle objective {cost:.12e}", flush=True)
        assert s.iterations == steps and s.final_cost < s.initial_cost
        assert math.isclose(cost, s.final_cost, rel_tol=1e-9)
        out[mv] = s
    s64, s32 = out[None], out[torch.float32]
    assert (s32.accepted_steps, s32.rejected_steps) == (s64.accepted_steps, s64.rejected_steps)
    assert abs(s32.final_cost - s64.final_cost) <= 1e-6 * s64.final_cost


def dense_reduced_matrix(J, diag, lam):
    """blockdiag(U + lam I) - sum over points p and observation pairs (i, j) of p of
    W_i (V_p + lam I)^-1 W_j^T at camera block (c_i, c_j), W_i = J_c,i^T J_p,i --
    the Schur complement written out pair by pair (vectorised over observations)."""
    g = J.groups[0]
    Jc, Jp = g.blocks
    cam, pt = g.indices
    nc = J.layout[0][0]
    W = np.einsum("nrc,nrp->ncp", Jc, Jp)
    Cinv = np.linalg.inv(diag[1] + lam * np.eye(3))
    order = np.argsort(pt, kind="stable")
    cam, pt, W = cam[order], pt[order], W[order]
    starts = np.concatenate(([0], np.flatnonzero(np.diff(pt)) + 1, [pt.size]))
    lens = np.diff(starts)
    M = np.zeros((nc * nc, 36))
    for L in np.unique(lens):
        s0 = starts[:-1][lens == L]
        ii = (s0[:, None, None] + np.arange(L)[None, :, None]).repeat(L, 2).reshape(-1)
        jj = (s0[:, None, None] + np.arange(L)[None, None, :]).repeat(L, 1).reshape(-1)
        blocks = np.einsum("ncp,npq,ndq->ncd", W[ii], Cinv[pt[ii]], W[jj]).reshape(-1, 36)
        key = cam[ii] * nc + cam[jj]
        for e in range(36):
            M[:, e] -= np.bincount(key, weights=blocks[:, e], minlength=nc * nc)
    M = M.reshape(nc, nc, 6, 6)
    M[np.arange(nc), np.arange(nc)] += diag[0] + lam * np.eye(6)
    return M.transpose(0, 2, 1, 3).reshape(6 * nc, 6 * nc)


def test_dense_schur_assembly_256(device):
    """c5 dense Schur: the 1,536 x 1,536 reduced matrix of 256 cameras / 20,000 points."""
    prob = generators.ba_dense_schur()
    print(f"dense schur assembly cams={prob.n_cameras} obs={prob.n_obs}", flush=True)
    P = ba_problem(prob, device)
    an = P._analyze()
    lin = an.linearize(P.parameters)
    diag = tuple(d.double() for d in lin.jacobian.diagonal_blocks())
    lam = 1e-3
    system = HipSchurPlan(an, "points", device).linearize(lin.jacobian, diag, dense=True).damped(lam)
    M = system.dense_matrix().cpu().numpy()
    ref = dense_reduced_matrix(to_np(lin.jacobian, an.layout), [d.cpu().numpy() for d in diag], lam)
    err = rel_err(np.tril(M), np.tril(ref))
    print(f"  rel err {err:.2e}", flush=True)
    assert err <= 1e-12
    # and the dense solve through the MFMA Cholesky equals numpy's solve of the same matrix
    b = np.random.default_rng(9).standard_normal(ref.shape[0])
    x = system.dense_solve(torch.tensor(b, device=device)).cpu().numpy()
    assert rel_err(x, np.linalg.solve(ref, b)) <= 1e-9


def test_pose_graph_10k(device):
    """c5 pose graph (10,000 poses, 10,284 between factors, started at the true poses
    perturbed by 0.01 per tangent coordinate), full-system PCG + block Jacobi in
    float64 and in mixed precision.  The chain's graph Laplacian has a condition
    number of order n^2, so block-Jacobi PCG (the reference's preconditioner) meets
    rtol 1e-6 within the iteration cap only in the first three LM iterations; over those
    iterations every PCG solve must converge, the two precisions must follow the
    same cost history within 1e-6, and each reported final cost must equal the
    objective recomputed on the CPU (torch float64) at the returned poses."""
    from torchls_amd.geometry import se3_inv, se3_log

    pg = generators.pose_graph(init_noise=0.01)
    steps = 3  # the PCG solves of the first three LM iterations meet rtol (418, 998, 2,371 iterations)
    print(f"pose graph poses={pg.n_poses} factors={pg.n_factors}, {steps} LM iterations", flush=True)
    out = {}
    for mv in (None, torch.float32):
        P = pose_graph_problem(pg, device)
        # mixed: float32 PCG vectors end at their first converged recurrence, without
        # restarts (their true residual's rounding floor, above rtol here, would fire the
        # drift test at every check); the float64 LM iteration corrects the step
        mixed = dict(stop_on_stagnation=True, max_refinements=0, drift_restarts=False) if mv is not None else {}
        cg = ConjugateGradientConfig(matvec_dtype=mv, max_iterations=3000, **mixed)
        vals, s = P.solve(return_summary=True, conjugate_gradient=cg,
                          termination=TerminationConfig(max_iterations=steps))
        print(f"  matvec {mv}: {s.termination_reason}, costs {[f'{c:.9e}' for c in s.cost_history]}, "
              f"pcg {s.linear_iterations_history}, |g| {s.gradient_norm:.2e}", flush=True)
        assert s.iterations == steps and s.rejected_steps == 0
        assert all(s.linear_converged_history) or mv is not None
        T = vals["poses"].double().cpu()
        Z = torch.tensor(pg.measurements)
        r = se3_log(se3_inv(Z) @ se3_inv(T[pg.i]) @ T[pg.j])
        cost = 0.5 * float(r.square().sum())
        assert math.isclose(cost, s.final_cost, rel_tol=1e-9)
        out[mv] = s
    s64, s32 = out[None], out[torch.float32]
    # mixed precision stores the blocks in float32 (the gradient is summed in float64 from
    # them: 6e-8 relative after the first step) and runs float32 PCG vectors, whose
    # recurrence on this system (condition ~1e8) takes about as many iterations as
    # float64's and leaves slightly different steps
    print(f"  cost rel diff per iteration {[abs(a - b) / b for a, b in zip(s32.cost_history, s64.cost_history)]}",
          flush=True)
    assert math.isclose(s32.cost_history[1], s64.cost_history[1], rel_tol=2e-7)
    assert np.allclose(s32.cost_history, s64.cost_history, rtol=1e-6, atol=0)
    assert sum(s32.linear_iterations_history) <= 1.1 * sum(s64.linear_iterations_history)
