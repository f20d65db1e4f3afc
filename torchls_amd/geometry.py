"""SE(3)/SO(3) helpers and the SE3Var manifold (tensor-only, torch.func compatible).

Conventions of the reference (geometry.py:40-57, 81-92, 130-137): rotation
vectors are axis-angle; the SE(3) tangent is ordered [translation(3),
rotation(3)]; poses are 4x4 matrices updated on the right, T <- T Exp(delta).
The exponential uses the closed-form Rodrigues coefficients
a = sin t / t, b = (1 - cos t) / t^2, c = (t - sin t) / t^3, switched to their
Taylor series for t^2 < 1e-4 so derivatives stay finite at zero.
"""

from __future__ import annotations

import torch
from torch import Tensor

from .variables import Var


def skew(v: Tensor) -> Tensor:
    x, y, z = v.unbind(-1)
    o = torch.zeros_like(x)
    return torch.stack((o, -z, y, z, o, -x, -y, x, o), dim=-1).reshape(*v.shape[:-1], 3, 3)


def _rodrigues(w: Tensor):
    t2 = (w * w).sum(dim=-1, keepdim=True)
    small = t2 < 1e-4
    safe = torch.where(small, torch.ones_like(t2), t2)
    t = safe.sqrt()
    s, c_ = torch.sin(t), torch.cos(t)
    a = torch.where(small, 1 - t2 / 6 + t2 * t2 / 120, s / t)
    b = torch.where(small, 0.5 - t2 / 24 + t2 * t2 / 720, (1 - c_) / safe)
    c = torch.where(small, 1.0 / 6 - t2 / 120 + t2 * t2 / 5040, (t - s) / (safe * t))
    return a[..., None], b[..., None], c[..., None]


def so3_exp(w: Tensor) -> Tensor:
    K = skew(w)
    a, b, _ = _rodrigues(w)
    return torch.eye(3, dtype=w.dtype, device=w.device) + a * K + b * (K @ K)


def se3_exp(xi: Tensor) -> Tensor:
    """[translation(3), rotation(3)] -> 4x4 SE(3) matrix."""
    v, w = xi[..., :3], xi[..., 3:]
    K = skew(w)
    a, b, c = _rodrigues(w)
    eye = torch.eye(3, dtype=xi.dtype, device=xi.device)
    KK = K @ K
    R = eye + a * K + b * KK
    V = eye + b * K + c * KK
    top = torch.cat((R, V @ v[..., None]), dim=-1)
    bottom = xi.new_tensor([0.0, 0.0, 0.0, 1.0]).expand(*xi.shape[:-1], 1, 4)
    return torch.cat((top, bottom), dim=-2)


def se3_retract_jvp(pose: Tensor, direction: Tensor) -> Tensor:
    """D_delta (T Exp(delta)) at delta = 0 applied to ``direction`` (storage space 4x4)."""
    R = pose[..., :3, :3]
    top = torch.cat((R @ skew(direction[..., 3:6]), R @ direction[..., :3, None]), dim=-1)
    return torch.cat((top, torch.zeros_like(top[..., :1, :])), dim=-2)


def so3_log(R: Tensor) -> Tensor:
    """Rotation matrix -> axis-angle via the unit quaternion (stable near 0 and pi)."""
    r00, r01, r02 = R[..., 0, 0], R[..., 0, 1], R[..., 0, 2]
    r10, r11, r12 = R[..., 1, 0], R[..., 1, 1], R[..., 1, 2]
    r20, r21, r22 = R[..., 2, 0], R[..., 2, 1], R[..., 2, 2]
    m = torch.stack((1 + r00 + r11 + r22, 1 + r00 - r11 - r22, 1 - r00 + r11 - r22,
                     1 - r00 - r11 + r22), -1).clamp_min(1e-12).sqrt()
    cand = torch.stack(
        (
            torch.stack((m[..., 0] ** 2, r21 - r12, r02 - r20, r10 - r01), -1),
            torch.stack((r21 - r12, m[..., 1] ** 2, r10 + r01, r02 + r20), -1),
            torch.stack((r02 - r20, r10 + r01, m[..., 2] ** 2, r12 + r21), -1),
            torch.stack((r10 - r01, r02 + r20, r12 + r21, m[..., 3] ** 2), -1),
        ),
        -2,
    ) / (2 * m[..., :, None])
    pick = m.argmax(-1)[..., None, None].expand(*m.shape[:-1], 1, 4)
    q = cand.gather(-2, pick).squeeze(-2)
    q = torch.where(q[..., :1] < 0, -q, q)
    n = (q[..., 1:] ** 2).sum(-1, keepdim=True).clamp_min(1e-24).sqrt()
    return q[..., 1:] * (2 * torch.atan2(n, q[..., :1]) / n)


def se3_log(T: Tensor) -> Tensor:
    w = so3_log(T[..., :3, :3])
    K = skew(w)
    t2 = (w * w).sum(-1, keepdim=True)
    small = t2 < 1e-4
    safe = torch.where(small, torch.ones_like(t2), t2)
    t = safe.sqrt()
    coef = torch.where(small, 1.0 / 12 + t2 / 720, (1 - t * torch.sin(t) / (2 * (1 - torch.cos(t)))) / safe)
    eye = torch.eye(3, dtype=T.dtype, device=T.device)
    Vinv = eye - 0.5 * K + coef[..., None] * (K @ K)
    return torch.cat(((Vinv @ T[..., :3, 3:4])[..., 0], w), -1)


def se3_inv(T: Tensor) -> Tensor:
    """Inverse of 4x4 SE(3) matrices: [R^T, -R^T t; 0, 1]."""
    Rt = T[..., :3, :3].transpose(-1, -2)
    top = torch.cat((Rt, -(Rt @ T[..., :3, 3:4])), dim=-1)
    return torch.cat((top, T[..., 3:4, :]), dim=-2)


def _se3_retract(pose: Tensor, delta: Tensor) -> Tensor:
    return pose @ se3_exp(delta)


class SE3Var(
    Var,
    default_factory=lambda: torch.eye(4),
    tangent_dim=6,
    retract_fn=_se3_retract,
    retract_jvp_fn=se3_retract_jvp,
):
    """SE(3) 4x4 poses; right update with a [translation(3), rotation(3)] tangent."""
