"""Seeded synthetic bundle-adjustment and pose-graph problems.

Everything is generated in float64 with numpy's PCG64 generator, so the same
seed yields bit-identical inputs on every machine.  Poses are world-to-camera
4x4 SE(3) matrices updated on the right with a [translation, rotation] tangent
(the convention of torchls ``SE3Var``, geometry.py:130-137); pixels are
normalized image coordinates (examples/named_bundle_adjustment.py:15-17).

The reference ships no generators for these workloads; the constructions below
follow the workload descriptions:

* cameras on a radius-5 circle in the z=0 plane, optical axis through the
  origin, image "down" axis along world -z;
* points uniform in [-1, 1]^3, each seen by k distinct random cameras;
* observations = exact projection + N(0, sigma^2), optional uniform outliers;
* initial poses T*Exp([dt; dw]) and initial points p + N(0, 0.05^2).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


# ---------------------------------------------------------------- Lie helpers
def skew_np(v: np.ndarray) -> np.ndarray:
    x, y, z = v[..., 0], v[..., 1], v[..., 2]
    o = np.zeros_like(x)
    return np.stack((o, -z, y, z, o, -x, -y, x, o), axis=-1).reshape(*v.shape[:-1], 3, 3)


def _so3_coeffs(w: np.ndarray):
    t2 = np.sum(w * w, axis=-1)
    t = np.sqrt(t2)
    small = t2 < 1e-8
    ts = np.where(small, 1.0, t)
    a = np.where(small, 1 - t2 / 6 + t2 * t2 / 120, np.sin(ts) / ts)
    b = np.where(small, 0.5 - t2 / 24 + t2 * t2 / 720, (1 - np.cos(ts)) / (ts * ts))
    c = np.where(small, 1.0 / 6 - t2 / 120 + t2 * t2 / 5040, (ts - np.sin(ts)) / (ts * ts * ts))
    return a[..., None, None], b[..., None, None], c[..., None, None]


def so3_exp_np(w: np.ndarray) -> np.ndarray:
    K = skew_np(w)
    a, b, _ = _so3_coeffs(w)
    return np.eye(3) + a * K + b * (K @ K)


def se3_exp_np(xi: np.ndarray) -> np.ndarray:
    """[translation(3), rotation(3)] -> 4x4, closed-form Rodrigues formulas."""
    v, w = xi[..., :3], xi[..., 3:]
    K = skew_np(w)
    a, b, c = _so3_coeffs(w)
    R = np.eye(3) + a * K + b * (K @ K)
    V = np.eye(3) + b * K + c * (K @ K)
    out = np.zeros(xi.shape[:-1] + (4, 4))
    out[..., :3, :3] = R
    out[..., :3, 3] = (V @ v[..., None])[..., 0]
    out[..., 3, 3] = 1.0
    return out


def so3_log_np(R: np.ndarray) -> np.ndarray:
    """Rotation matrix -> axis-angle, via the unit quaternion (stable near 0 and pi)."""
    tr = R[..., 0, 0] + R[..., 1, 1] + R[..., 2, 2]
    mags = np.stack(
        (1 + tr, 1 + 2 * R[..., 0, 0] - tr, 1 + 2 * R[..., 1, 1] - tr, 1 + 2 * R[..., 2, 2] - tr),
        axis=-1,
    )
    k = np.argmax(mags, axis=-1)
    q = np.empty(R.shape[:-2] + (4,))
    m = np.sqrt(np.maximum(np.take_along_axis(mags, k[..., None], -1)[..., 0], 1e-300))
    s = 0.5 / m
    w_ = np.where(k == 0, 0.5 * m, 0)
    x_ = np.where(k == 1, 0.5 * m, 0)
    y_ = np.where(k == 2, 0.5 * m, 0)
    z_ = np.where(k == 3, 0.5 * m, 0)
    r21, r12 = R[..., 2, 1], R[..., 1, 2]
    r02, r20 = R[..., 0, 2], R[..., 2, 0]
    r10, r01 = R[..., 1, 0], R[..., 0, 1]
    q[..., 0] = np.where(k == 0, w_, (np.choose(k, [r21 - r12, r21 - r12, r02 - r20, r10 - r01]) * s))
    q[..., 1] = np.where(k == 1, x_, np.choose(k, [r21 - r12, 0 * r21, r10 + r01, r02 + r20]) * s)
    q[..., 2] = np.where(k == 2, y_, np.choose(k, [r02 - r20, r10 + r01, 0 * r21, r12 + r21]) * s)
    q[..., 3] = np.where(k == 3, z_, np.choose(k, [r10 - r01, r02 + r20, r12 + r21, 0 * r21]) * s)
    q = np.where(q[..., :1] < 0, -q, q)
    vec = q[..., 1:]
    n = np.sqrt(np.sum(vec * vec, axis=-1))
    angle = 2 * np.arctan2(n, q[..., 0])
    scale = np.where(n < 1e-12, 2.0 / np.maximum(q[..., 0], 1e-300), angle / np.maximum(n, 1e-300))
    return vec * scale[..., None]


def se3_log_np(T: np.ndarray) -> np.ndarray:
    w = so3_log_np(T[..., :3, :3])
    K = skew_np(w)
    t2 = np.sum(w * w, axis=-1)
    t = np.sqrt(t2)
    small = t2 < 1e-8
    ts = np.where(small, 1.0, t)
    # V^-1 = I - K/2 + (1/t^2)(1 - (t sin t)/(2(1-cos t))) K^2
    coef = np.where(
        small, 1.0 / 12 + t2 / 720, (1 - ts * np.sin(ts) / (2 * (1 - np.cos(ts)))) / (ts * ts)
    )
    Vinv = np.eye(3) - 0.5 * K + coef[..., None, None] * (K @ K)
    v = (Vinv @ T[..., :3, 3:4])[..., 0]
    return np.concatenate((v, w), axis=-1)


def se3_inv_np(T: np.ndarray) -> np.ndarray:
    out = np.zeros_like(T)
    Rt = np.swapaxes(T[..., :3, :3], -1, -2)
    out[..., :3, :3] = Rt
    out[..., :3, 3] = -(Rt @ T[..., :3, 3:4])[..., 0]
    out[..., 3, 3] = 1.0
    return out


def look_at_world_to_camera(centers: np.ndarray, target=np.zeros(3), up=np.array([0.0, 0.0, 1.0])):
    f = target - centers
    f = f / np.linalg.norm(f, axis=-1, keepdims=True)
    right = np.cross(f, up)
    right = right / np.linalg.norm(right, axis=-1, keepdims=True)
    down = np.cross(f, right)
    R = np.stack((right, down, f), axis=-2)  # rows: camera axes in world coordinates
    T = np.zeros(centers.shape[:-1] + (4, 4))
    T[..., :3, :3] = R
    T[..., :3, 3] = -(R @ centers[..., None])[..., 0]
    T[..., 3, 3] = 1.0
    return T


def project_np(T: np.ndarray, p: np.ndarray) -> np.ndarray:
    local = (T[..., :3, :3] @ p[..., None])[..., 0] + T[..., :3, 3]
    return local[..., :2] / local[..., 2:3]


# ----------------------------------------------------------- visibility sampling
def _distinct_cameras(rng, point_of_obs: np.ndarray, n_cameras: int) -> np.ndarray:
    """Uniform camera per observation, redrawn until distinct within each point."""
    cams = rng.integers(0, n_cameras, size=point_of_obs.size)
    while True:
        key = point_of_obs.astype(np.int64) * n_cameras + cams
        order = np.argsort(key, kind="stable")
        dup = np.zeros(key.size, dtype=bool)
        dup[order[1:]] = key[order[1:]] == key[order[:-1]]
        count = int(dup.sum())
        if count == 0:
            return cams
        cams[dup] = rng.integers(0, n_cameras, size=count)


@dataclass
class BundleAdjustment:
    """A seeded synthetic bundle-adjustment problem (float64 numpy arrays)."""

    seed: int
    cameras_true: np.ndarray  # [n_c, 4, 4] world-to-camera
    points_true: np.ndarray  # [n_p, 3]
    cameras_init: np.ndarray
    points_init: np.ndarray
    cam_idx: np.ndarray  # [N] int64
    pt_idx: np.ndarray  # [N] int64
    observations: np.ndarray  # [N, 2]
    outliers: np.ndarray  # [N] bool
    huber_delta: float | None = 1e-2
    meta: dict = field(default_factory=dict)

    @property
    def n_cameras(self):
        return self.cameras_true.shape[0]

    @property
    def n_points(self):
        return self.points_true.shape[0]

    @property
    def n_obs(self):
        return self.cam_idx.size

    @property
    def tangent_dim(self):
        return 6 * self.n_cameras + 3 * self.n_points


def bundle_adjustment(
    n_cameras: int,
    n_points: int,
    obs_per_point: int | str = 8,
    *,
    seed: int = 1,
    noise: float = 1e-3,
    outlier_fraction: float = 0.0,
    rot_perturb: float = 0.01,
    trans_perturb: float = 0.05,
    point_perturb: float = 0.05,
    radius: float = 5.0,
    huber_delta: float | None = 1e-2,
    shuffle: bool = True,
) -> BundleAdjustment:
    """Cameras on a circle looking at the origin; points in [-1,1]^3.

    ``obs_per_point`` is an integer k, or ``"geometric"`` for k ~ Geometric(0.1)
    on {1, 2, ...} (mean 10); k is capped at the camera count.  Observations are
    returned in a seeded random order unless ``shuffle`` is False.
    """
    rng = np.random.default_rng(seed)
    theta = 2 * math.pi * np.arange(n_cameras) / n_cameras
    centers = np.stack((radius * np.cos(theta), radius * np.sin(theta), np.zeros(n_cameras)), -1)
    cams = look_at_world_to_camera(centers)
    points = rng.uniform(-1.0, 1.0, size=(n_points, 3))
    if obs_per_point == "geometric":
        k = rng.geometric(0.1, size=n_points)
    else:
        k = np.full(n_points, int(obs_per_point))
    k = np.minimum(k, n_cameras)
    pt_idx = np.repeat(np.arange(n_points), k)
    cam_idx = _distinct_cameras(rng, pt_idx, n_cameras)
    uv = project_np(cams[cam_idx], points[pt_idx])
    uv = uv + rng.normal(0.0, noise, size=uv.shape)
    outliers = np.zeros(uv.shape[0], dtype=bool)
    if outlier_fraction > 0:
        outliers = rng.random(uv.shape[0]) < outlier_fraction
        uv[outliers] = rng.uniform(-1.0, 1.0, size=(int(outliers.sum()), 2))
    if shuffle:
        perm = rng.permutation(uv.shape[0])
        cam_idx, pt_idx, uv, outliers = cam_idx[perm], pt_idx[perm], uv[perm], outliers[perm]
    xi = np.concatenate(
        (
            rng.normal(0.0, trans_perturb, size=(n_cameras, 3)),
            rng.normal(0.0, rot_perturb, size=(n_cameras, 3)),
        ),
        axis=-1,
    )
    cams0 = cams @ se3_exp_np(xi)
    pts0 = points + rng.normal(0.0, point_perturb, size=points.shape)
    return BundleAdjustment(
        seed,
        cams,
        points,
        cams0,
        pts0,
        cam_idx.astype(np.int64),
        pt_idx.astype(np.int64),
        uv,
        outliers,
        huber_delta,
        dict(noise=noise, outlier_fraction=outlier_fraction, obs_per_point=obs_per_point),
    )


# Named workloads (sizes of the five benchmark configurations).
def ba_small(seed: int = 1) -> BundleAdjustment:
    """16 cameras, 500 points, 8 observations per point (4,000 residual blocks)."""
    return bundle_adjustment(16, 500, 8, seed=seed)


def ba_medium(seed: int = 2) -> BundleAdjustment:
    """200 cameras, 50,000 points, 10 observations per point (500,000 residual blocks)."""
    return bundle_adjustment(200, 50_000, 10, seed=seed)


def ba_large(seed: int = 3) -> BundleAdjustment:
    """2,000 cameras, 500,000 points, geometric(0.1) views per point, 5% outliers."""
    return bundle_adjustment(2_000, 500_000, "geometric", seed=seed, outlier_fraction=0.05)


def ba_differentiable(seed: int = 4) -> BundleAdjustment:
    """500 cameras, 100,000 points, 10 observations per point (1,000,000 residual blocks)."""
    return bundle_adjustment(500, 100_000, 10, seed=seed)


def ba_dense_schur(seed: int = 5, n_points: int = 20_000) -> BundleAdjustment:
    """256 cameras (a 1,536-dimensional reduced camera system)."""
    return bundle_adjustment(256, n_points, 10, seed=seed)


@dataclass
class PoseGraph:
    seed: int
    poses_true: np.ndarray  # [n, 4, 4] (world-from-body)
    poses_init: np.ndarray
    i: np.ndarray  # [F] int64
    j: np.ndarray  # [F] int64
    measurements: np.ndarray  # [F, 4, 4]  Z_ij ~ T_i^-1 T_j
    n_loop: int = 0

    @property
    def n_poses(self):
        return self.poses_true.shape[0]

    @property
    def n_factors(self):
        return self.i.size


def pose_graph(
    n_poses: int = 10_000,
    *,
    seed: int = 5,
    step: float = 0.5,
    yaw_noise: float = 0.1,
    loop_fraction: float = 0.03,
    loop_radius: float = 2.0,
    noise: float = 0.01,
    init_noise: float | None = None,
) -> PoseGraph:
    """Planar random walk in SE(3); odometry plus loop closures within ``loop_radius``.

    Loop closures are sampled among pose pairs (|i-j| > 5) whose positions lie within
    ``loop_radius``; their number is ``loop_fraction * n_poses`` (fewer if not enough
    candidate pairs exist).  Each measurement is the true relative pose times
    Exp(N(0, noise^2 I_6)).  The initial guess integrates the noisy odometry, or,
    with ``init_noise``, is each true pose times Exp(N(0, init_noise^2 I_6)) (pose 0
    exact): a start inside the basin of the optimum, from which LM converges.
    """
    rng = np.random.default_rng(seed)
    yaw = np.cumsum(np.concatenate(([0.0], rng.normal(0.0, yaw_noise, size=n_poses - 1))))
    heading = np.stack((np.cos(yaw), np.sin(yaw), np.zeros(n_poses)), -1)
    pos = np.concatenate((np.zeros((1, 3)), np.cumsum(step * heading[:-1], axis=0)))
    T = np.zeros((n_poses, 4, 4))
    T[:, :3, :3] = so3_exp_np(np.stack((np.zeros(n_poses), np.zeros(n_poses), yaw), -1))
    T[:, :3, 3] = pos
    T[:, 3, 3] = 1
    odo_i = np.arange(n_poses - 1)
    odo_j = odo_i + 1
    # loop closure candidates: spatial hashing on a grid of cell size loop_radius
    want = int(round(loop_fraction * n_poses))
    cell = np.floor(pos[:, :2] / loop_radius).astype(np.int64)
    buckets: dict = {}
    for idx, key in enumerate(map(tuple, cell)):
        buckets.setdefault(key, []).append(idx)
    cand_i, cand_j = [], []
    for (cx, cy), members in buckets.items():
        neigh = []
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                neigh.extend(buckets.get((cx + dx, cy + dy), ()))
        neigh = np.asarray(neigh)
        for a in members:
            close = neigh[(neigh > a + 5)]
            if close.size:
                d = np.linalg.norm(pos[close] - pos[a], axis=-1)
                sel = close[d <= loop_radius]
                cand_i.extend([a] * sel.size)
                cand_j.extend(sel.tolist())
    cand_i, cand_j = np.asarray(cand_i, dtype=np.int64), np.asarray(cand_j, dtype=np.int64)
    n_loop = min(want, cand_i.size)
    pick = rng.choice(cand_i.size, size=n_loop, replace=False) if n_loop else np.zeros(0, np.int64)
    fi = np.concatenate((odo_i, cand_i[pick]))
    fj = np.concatenate((odo_j, cand_j[pick]))
    rel = se3_inv_np(T[fi]) @ T[fj]
    Z = rel @ se3_exp_np(rng.normal(0.0, noise, size=(fi.size, 6)))
    # initial guess: chain the noisy odometry
    if init_noise is None:
        init = np.zeros_like(T)
        init[0] = T[0]
        for k in range(n_poses - 1):
            init[k + 1] = init[k] @ Z[k]
    else:
        init = T @ se3_exp_np(rng.normal(0.0, init_noise, size=(n_poses, 6)))
        init[0] = T[0]
    return PoseGraph(seed, T, init, fi, fj, Z, n_loop)
