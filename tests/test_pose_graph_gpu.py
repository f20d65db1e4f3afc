"""Pose-graph LM (SE(3) between factors, aliased 6x6 blocks) on the HIP backend."""

import numpy as np
import pytest
import torch

from torchls_amd import ConjugateGradientConfig, TerminationConfig, generators
from torchls_amd.workloads import pose_graph_problem

pytestmark = pytest.mark.gpu


def _solve(pg, device, linear_solver, mv_dtype):
    # a 300-pose chain is ill-conditioned (its graph Laplacian's condition grows with
    # the square of its length): block-Jacobi PCG needs a few hundred iterations per step
    P = pose_graph_problem(pg, device)
    vals, s = P.solve(return_summary=True, linear_solver=linear_solver,
                      conjugate_gradient=ConjugateGradientConfig(matvec_dtype=mv_dtype, max_iterations=2000),
                      termination=TerminationConfig(max_iterations=30))
    return vals["poses"].cpu().numpy(), s


def _relative(T, i, j):
    Ti, Tj = T[i], T[j]
    Rt = np.swapaxes(Ti[:, :3, :3], -1, -2)
    return Rt @ Tj[:, :3, :3], (Rt @ (Tj[:, :3, 3] - Ti[:, :3, 3])[..., None])[..., 0]


def test_full_system_pcg_matches_dense_qr(device):
    """Full-system PCG + block-Jacobi (float64 and float32 matvec) reaches the dense-QR LM's
    optimum on a 300-pose graph with loop closures (gauge-free comparison: relative poses)."""
    pg = generators.pose_graph(300, seed=3, loop_fraction=0.05, loop_radius=3.0)
    print(f"pose graph poses={pg.n_poses} factors={pg.n_factors} loops={pg.n_loop}", flush=True)
    T_qr, s_qr = _solve(pg, device, "dense_qr", None)
    assert s_qr.final_cost < 0.2 * s_qr.initial_cost  # odometry is near the noise level; loops are not
    for mv in (None, torch.float32):
        T, s = _solve(pg, device, "conjugate_gradient", mv)
        print(f"  matvec {mv}: final {s.final_cost:.12e} vs QR {s_qr.final_cost:.12e}; LM {s.iterations}, "
              f"pcg {s.linear_iterations_history}", flush=True)
        assert abs(s.final_cost - s_qr.final_cost) <= 1e-6 * s_qr.final_cost
        R, t = _relative(T, pg.i, pg.j)
        Rq, tq = _relative(T_qr, pg.i, pg.j)
        assert np.abs(R - Rq).max() <= 1e-5 and np.abs(t - tq).max() <= 1e-5


def _plain_between(pose_i, pose_j, measurement):
    """The same residual without the HIP marker: linearized by torch.func (vmap of jacrev)."""
    from torchls_amd.geometry import se3_inv, se3_log

    return se3_log(se3_inv(measurement) @ se3_inv(pose_i) @ pose_j)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-11), (torch.float32, 2e-5)])
def test_hip_between_linearization_matches_autodiff(device, dtype, tol):
    """tlh_between_linearize (dual numbers) against torch.func's Jacobian of the same logarithm,
    at the noisy initial poses (loop-closure residuals of 0.1-1 rad) and at rotations near pi."""
    from torchls_amd import Problem, SE3Var
    from torchls_amd.generators import se3_exp_np
    from torchls_amd.residuals import between

    pg = generators.pose_graph(300, seed=3, loop_fraction=0.05, loop_radius=3.0)
    rng = np.random.default_rng(4)
    poses = pg.poses_init.copy()
    poses[:20] = poses[:20] @ se3_exp_np(np.concatenate((rng.normal(0, 1, (20, 3)), rng.normal(0, 1.2, (20, 3))), -1))
    lins = []
    for fn in (between, _plain_between):
        P = Problem({"poses": torch.tensor(poses, dtype=dtype, device=device)}, manifolds={"poses": SE3Var})
        P.add_residual(fn, ("poses", torch.tensor(pg.i)), ("poses", torch.tensor(pg.j)),
                       torch.tensor(pg.measurements, dtype=dtype, device=device))
        lins.append(P.linearize())
    (a, b) = lins
    print(f"between linearization dtype={dtype}", flush=True)
    assert float(a.nonfinite) == 0
    ra, rb = a.residual.double().cpu(), b.residual.double().cpu()
    assert (ra - rb).abs().max() <= tol * max(1.0, rb.abs().max())
    for ga, gb in zip(a.jacobian.groups[0].blocks, b.jacobian.groups[0].blocks):
        ga, gb = ga.double().cpu(), gb.double().cpu()
        assert (ga - gb).norm() <= tol * gb.norm(), float((ga - gb).norm() / gb.norm())
