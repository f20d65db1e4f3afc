"""Residual templates of the benchmark problems, with analytic Jacobians.

Single-observation functions work with ``Problem.add_residual`` (vmapped like
any user residual); the ``*_batched`` helpers linearize a whole batch at once
on the device and are what the synthetic benchmarks use.

Reprojection follows examples/named_bundle_adjustment.py:15-17 with a
world-to-camera SE(3) pose updated on the right, tangent [v; w]
(geometry.py:130-137):  d(Rp+t)/dv = R, d(Rp+t)/dw = -R[p]x, d(Rp+t)/dp = R.
"""

from __future__ import annotations

import torch
from torch import Tensor

from .geometry import se3_inv, se3_log, skew


def reprojection(camera: Tensor, point: Tensor, observation: Tensor) -> Tensor:
    local = camera[:3, :3] @ point + camera[:3, 3]
    return local[:2] / local[2] - observation


def reprojection_jacobian(camera: Tensor, point: Tensor, observation: Tensor):
    """(d r / d camera-tangent [2,6], d r / d point [2,3]) for one observation."""
    R = camera[:3, :3]
    local = R @ point + camera[:3, 3]
    x, y, z = local[0], local[1], local[2]
    zero = torch.zeros_like(z)
    D = torch.stack((1 / z, zero, -x / (z * z), zero, 1 / z, -y / (z * z))).reshape(2, 3)
    DR = D @ R
    return torch.cat((DR, -(DR @ skew(point))), dim=-1), DR


def reprojection_batched(cameras: Tensor, points: Tensor, cam_idx: Tensor, pt_idx: Tensor, uv: Tensor):
    """Residuals [N,2] and blocks [N,2,6], [N,2,3] for all observations at once."""
    R = cameras[cam_idx, :3, :3]
    p = points[pt_idx]
    local = (R @ p[..., None])[..., 0] + cameras[cam_idx, :3, 3]
    x, y, z = local.unbind(-1)
    res = torch.stack((x / z, y / z), -1) - uv
    inv = 1 / z
    D = torch.zeros(z.shape + (2, 3), dtype=z.dtype, device=z.device)
    D[:, 0, 0] = inv
    D[:, 0, 2] = -x * inv * inv
    D[:, 1, 1] = inv
    D[:, 1, 2] = -y * inv * inv
    DR = D @ R
    Jc = torch.cat((DR, -(DR @ skew(p))), dim=-1)
    return res, Jc, DR


def _reprojection_rows(cams: Tensor, pts: Tensor, uv: Tensor):
    R = cams[:, :3, :3]
    local = (R @ pts[..., None])[..., 0] + cams[:, :3, 3]
    x, y, z = local.unbind(-1)
    return R, x, y, z, torch.stack((x / z, y / z), -1) - uv


def _batched_residual(cams: Tensor, pts: Tensor, uv: Tensor) -> Tensor:
    return _reprojection_rows(cams, pts, uv)[-1]


def _batched_linearize(cams: Tensor, pts: Tensor, uv: Tensor):
    R, x, y, z, res = _reprojection_rows(cams, pts, uv)
    inv = 1 / z
    D = torch.zeros(z.shape + (2, 3), dtype=z.dtype, device=z.device)
    D[:, 0, 0] = inv
    D[:, 0, 2] = -x * inv * inv
    D[:, 1, 1] = inv
    D[:, 1, 2] = -y * inv * inv
    DR = D @ R
    return res, (torch.cat((DR, -(DR @ skew(pts))), dim=-1), DR)


# Batched evaluators used by Problem instead of vmap for this residual; with
# SE(3) poses and Euclidean points Problem runs the HIP kernels
# tlh_reprojection_linearize / tlh_reprojection_trial instead (hip_ops).
reprojection.batched_residual = _batched_residual
reprojection.batched_linearize = _batched_linearize
reprojection.hip_residual = "reprojection"


def between(pose_i: Tensor, pose_j: Tensor, measurement: Tensor) -> Tensor:
    """SE(3) between factor of the pose-graph workload: Log(Z^-1 T_i^-1 T_j), 6 rows,
    [translation, rotation] like the tangent.  The reference's geometry has no SE(3)
    log; the logarithm here (geometry.se3_log) is the inverse of its exponential
    (geometry.py:40-57).  Linearized by torch.func like any user residual, giving the
    two aliased 6x6 blocks of a pose-graph edge."""
    return se3_log(se3_inv(measurement) @ se3_inv(pose_i) @ pose_j)


def _between_batched(pose_i: Tensor, pose_j: Tensor, measurement: Tensor) -> Tensor:
    return se3_log(se3_inv(measurement) @ se3_inv(pose_i) @ pose_j)


between.batched_residual = _between_batched
# with one SE(3) group, no loss and no weights Problem linearizes it with the HIP
# kernel tlh_between_linearize (forward-mode dual numbers, exact 6x12 blocks)
between.hip_residual = "between"
