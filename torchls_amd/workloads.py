"""Device-side construction of the synthetic benchmark linearizations."""

from __future__ import annotations

import torch

from .generators import BundleAdjustment, PoseGraph
from .residuals import reprojection_batched
from .sparse import BlockJacobian, JacobianGroup, VariableLayout


class _Tangent:
    def __init__(self, d: int):
        self.tangent_dim = d


def ba_layout(n_cameras: int, n_points: int):
    return (
        VariableLayout(_Tangent(6), torch.arange(n_cameras), 0, "cams"),
        VariableLayout(_Tangent(3), torch.arange(n_points), 6 * n_cameras, "points"),
    )


def huber_scale(res: torch.Tensor, delta: float | None) -> torch.Tensor:
    """sqrt(rho'(||r||^2)) of the block Huber loss (losses.py:71-78)."""
    s = res.double().square().sum(-1)
    if delta is None:
        return torch.ones_like(s)
    norm = s.clamp_min(torch.finfo(torch.float64).tiny).sqrt()
    return torch.where(s > delta * delta, delta / norm, torch.ones_like(norm)).sqrt()


def ba_linearization(prob: BundleAdjustment, device, dtype=torch.float64, *, at_truth: bool = False):
    """Huber-weighted residual [N*2] and BlockJacobian of a generated BA problem."""
    cams = torch.tensor(prob.cameras_true if at_truth else prob.cameras_init, dtype=dtype, device=device)
    pts = torch.tensor(prob.points_true if at_truth else prob.points_init, dtype=dtype, device=device)
    ci = torch.tensor(prob.cam_idx, device=device)
    pi = torch.tensor(prob.pt_idx, device=device)
    uv = torch.tensor(prob.observations, dtype=dtype, device=device)
    res, Jc, Jp = reprojection_batched(cams, pts, ci, pi, uv)
    w = huber_scale(res, prob.huber_delta).to(dtype)
    res = res * w[:, None]
    Jc = Jc * w[:, None, None]
    Jp = Jp * w[:, None, None]
    layout = ba_layout(prob.n_cameras, prob.n_points)
    ar6 = torch.arange(6, device=device)
    ar3 = torch.arange(3, device=device)
    cols = (ci[:, None] * 6 + ar6, pi[:, None] * 3 + 6 * prob.n_cameras + ar3)
    group = JacobianGroup((Jc.contiguous(), Jp.contiguous()), cols, (0, 1), (ci, pi))
    J = BlockJacobian((group,), layout, (2 * prob.n_obs, prob.tangent_dim))
    return res.reshape(-1), J


def normal_matvec_bytes(n_obs: int, rows: int, dims: tuple, n: int, esize: int) -> int:
    """Algorithmic HBM bytes of one normal matvec: blocks + two int32 ids per
    observation, plus reading x and writing y once."""
    return n_obs * (rows * sum(dims) * esize + 4 * len(dims)) + 2 * n * esize


def pose_graph_problem(pg: PoseGraph, device, dtype=torch.float64):
    """Problem of a generated pose graph: one SE(3) variable per pose and one
    between factor Log(Z^-1 T_i^-1 T_j) per odometry edge and loop closure."""
    from .geometry import SE3Var
    from .problem import Problem
    from .residuals import between

    P = Problem({"poses": torch.tensor(pg.poses_init, dtype=dtype, device=device)}, manifolds={"poses": SE3Var})
    P.add_residual(between, ("poses", torch.tensor(pg.i)), ("poses", torch.tensor(pg.j)),
                   torch.tensor(pg.measurements, dtype=dtype, device=device))
    return P
