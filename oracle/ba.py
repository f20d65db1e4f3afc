"""Analytic float64 linearization of the benchmark residuals (CPU oracle).

Reprojection (examples/named_bundle_adjustment.py:15-17):
    r(T, p) = (R p + t)_xy / (R p + t)_z - uv,   T = [R t; 0 1] world-to-camera.
Tangent convention of SE3Var (geometry.py:130-137, 81-92): T <- T Exp([v; w]);
the linearization is taken at delta = 0 of the affine chart
T + D Exp(0)[delta] (linearization.py:67-79), whose first derivative equals
that of the exact retraction, so

    d(R p + t)/dv = R,   d(R p + t)/dw = -R [p]x,   d(R p + t)/dp = R.

Robust loss: HuberLoss on the block norm scales the raw residual and both
blocks by sqrt(rho'(s)) (losses.py:88-90, problem.py:177-181 semantics).
"""

from __future__ import annotations

import numpy as np

from torchls_amd.generators import skew_np


def reprojection(cams: np.ndarray, pts: np.ndarray, cam_idx, pt_idx, uv):
    """Raw residuals [N, 2] and Jacobian blocks [N, 2, 6], [N, 2, 3]."""
    R = cams[cam_idx, :3, :3]
    t = cams[cam_idx, :3, 3]
    p = pts[pt_idx]
    local = (R @ p[..., None])[..., 0] + t
    x, y, z = local[:, 0], local[:, 1], local[:, 2]
    res = np.stack((x / z, y / z), -1) - uv
    D = np.zeros((len(z), 2, 3))
    D[:, 0, 0] = 1 / z
    D[:, 0, 2] = -x / (z * z)
    D[:, 1, 1] = 1 / z
    D[:, 1, 2] = -y / (z * z)
    DR = D @ R
    Jc = np.concatenate((DR, -(DR @ skew_np(p))), axis=-1)
    return res, Jc, DR


def huber_weights(res: np.ndarray, delta: float | None):
    """sqrt(rho'(s)) of the block Huber loss (losses.py:71-78) and the objective terms."""
    s = np.sum(res * res, axis=-1)
    if delta is None:
        return np.ones_like(s), s
    norm = np.sqrt(np.maximum(s, np.finfo(np.float64).tiny))
    out = s > delta * delta
    first = np.where(out, delta / norm, 1.0)
    value = np.where(out, delta * (2 * norm - delta), s)
    return np.sqrt(first), value


def linearize_ba(cams, pts, cam_idx, pt_idx, uv, delta=None):
    """Model residual, blocks (Huber-weighted), raw residual and 0.5*sum(rho)."""
    res, Jc, Jp = reprojection(cams, pts, cam_idx, pt_idx, uv)
    w, value = huber_weights(res, delta)
    return res * w[:, None], Jc * w[:, None, None], Jp * w[:, None, None], res, 0.5 * float(value.sum())
