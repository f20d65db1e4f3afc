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


# ----------------------------------------------------------------- problem object
from torchls_amd.generators import se3_exp_np  # noqa: E402

from .sparse import NpGroup, NpJacobian  # noqa: E402


def huber_value(s, delta):
    norm = np.sqrt(np.maximum(s, np.finfo(np.float64).tiny))
    return np.where(s > delta * delta, delta * (2 * norm - delta), s)


def huber_reduction(before, after, diff, delta):
    """Divided difference of the Huber loss (losses.py:95-107)."""
    tiny = np.finfo(np.float64).tiny
    a, b = np.sqrt(before), np.sqrt(after)
    low, high = np.minimum(a, b), np.maximum(a, b)
    short = np.maximum(delta - low, 0)
    crossing = (2 * delta - short * (short / np.maximum(high - low, tiny))) / np.maximum(a + b, tiny)
    coef = np.where(high <= delta, 1.0, np.where(low >= delta, 2 * delta / np.maximum(a + b, tiny), crossing))
    return diff * coef


class BAProblem:
    """Reprojection BA (+ optional gauge priors on camera 0 / point 0), float64.

    vals = (cams [n_c,4,4], pts [n_p,3]).  Mirrors the objective/linearization
    semantics of problem.py:20-42, 301-368 with HuberLoss (losses.py:56-107).
    """

    def __init__(self, cam_idx, pt_idx, uv, n_cams, n_pts, huber_delta=None, priors=None):
        self.cam_idx, self.pt_idx, self.uv = cam_idx, pt_idx, uv
        self.delta = huber_delta
        self.layout = [(n_cams, 6, 0), (n_pts, 3, 6 * n_cams)]
        self.priors = priors  # None or (cam0_ref [4,4], pt0_ref [3])
        self.has_loss = huber_delta is not None

    def _prior_terms(self, vals):
        cams, pts = vals
        Tref, pref = self.priors
        T = cams[0]
        rc = (T[:3, :] - Tref[:3, :]).reshape(-1)
        R = T[:3, :3]
        Jc = np.zeros((12, 6))
        for k in range(3):  # rotation part: R [e_k]x
            e = np.zeros(3)
            e[k] = 1
            from torchls_amd.generators import skew_np

            Jc[:, 3 + k] = np.concatenate((R @ skew_np(e), np.zeros((3, 1))), axis=1).reshape(-1)
            Jc[:, k] = np.concatenate((np.zeros((3, 3)), R[:, k : k + 1]), axis=1).reshape(-1)
        rp = pts[0] - pref
        return rc, Jc, rp, np.eye(3)

    def raw(self, vals):
        cams, pts = vals
        res, _, _ = reprojection(cams, pts, self.cam_idx, self.pt_idx, self.uv)
        return res

    def linearize(self, vals):
        cams, pts = vals
        res, Jc, Jp = reprojection(cams, pts, self.cam_idx, self.pt_idx, self.uv)
        raw = res
        s = np.sum(res * res, -1)
        if self.has_loss:
            w = np.sqrt(np.where(s > self.delta**2, self.delta / np.sqrt(np.maximum(s, np.finfo(float).tiny)), 1.0))
            res, Jc, Jp = res * w[:, None], Jc * w[:, None, None], Jp * w[:, None, None]
        groups = [NpGroup([Jc, Jp], [0, 1], [self.cam_idx, self.pt_idx])]
        residual = [res.reshape(-1)]
        raws = [raw.reshape(-1)]
        obj = float(np.sum(huber_value(s, self.delta))) if self.has_loss else float(np.sum(s))
        if self.priors is not None:
            rc, Jpc, rp, Jpp = self._prior_terms(vals)
            groups.append(NpGroup([Jpc[None]], [0], [np.zeros(1, np.int64)]))
            groups.append(NpGroup([Jpp[None]], [1], [np.zeros(1, np.int64)]))
            residual += [rc, rp]
            raws += [rc, rp]
            obj += float(rc @ rc + rp @ rp)
        J = NpJacobian(groups, self.layout)
        r = np.concatenate(residual)
        return dict(residual=r, J=J, raw=np.concatenate(raws) if self.has_loss else None,
                    objective=0.5 * obj if self.has_loss else None)

    def cost(self, vals):
        s = np.sum(self.raw(vals) ** 2, -1)
        c = float(np.sum(huber_value(s, self.delta))) if self.has_loss else float(np.sum(s))
        if self.priors is not None:
            rc, _, rp, _ = self._prior_terms(vals)
            c += float(rc @ rc + rp @ rp)
        return 0.5 * c

    def evaluate_trial(self, vals, lin):
        new = self.raw(vals)
        before = lin["raw"] if self.has_loss else lin["residual"]
        n = new.shape[0]
        old = before[: 2 * n].reshape(n, 2)
        so, sn = np.sum(old * old, -1), np.sum(new * new, -1)
        diff = np.sum((old - new) * (old + new), -1)
        if self.has_loss:
            obj = float(np.sum(huber_value(sn, self.delta)))
            red = float(np.sum(huber_reduction(so, sn, diff, self.delta)))
        else:
            obj, red = float(np.sum(sn)), float(np.sum(diff))
        if self.priors is not None:
            rc, _, rp, _ = self._prior_terms(vals)
            pn = np.concatenate((rc, rp))
            po = before[2 * n :]
            obj += float(pn @ pn)
            red += float((po - pn) @ (po + pn))
        return 0.5 * obj, 0.5 * red

    def retract(self, vals, delta):
        cams, pts = vals
        nc = cams.shape[0]
        dc = delta[: 6 * nc].reshape(nc, 6)
        dp = delta[6 * nc :].reshape(-1, 3)
        return cams @ se3_exp_np(dc), pts + dp
