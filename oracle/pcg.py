"""float64 PCG restatements (CPU oracle).

``pcg_reference`` follows torchls solvers.pcg (solvers.py:167-256): textbook
PCG from x0 = 0 with true-residual checks when the recurrence norm reaches the
threshold or every ``residual_check_interval`` iterations, and adaptive or
periodic restarts.

``pcg_batched`` follows warp_pcg._run (warp_pcg.py:110-214), the control flow
the HIP backend reproduces: device-side status word, replays in batches up to
the next check multiple, status 1 set from ||r||^2 <= thr^2 after the direction
update, restart on status 1 / periodic / drift.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class Result:
    solution: np.ndarray
    iterations: int
    residual_norm: float
    converged: bool
    reason: str
    restarts: int = 0
    residual_checks: int = 0
    rhs_norm: float = 0.0


@dataclass
class Config:
    rtol: float = 1e-6
    atol: float = 1e-10
    max_iterations: int = 200
    residual_check_interval: int = 50
    restart_strategy: str = "adaptive"
    residual_drift_tolerance: float = 0.1
    graph_batch_iterations: int = 16
    # the HIP backend's optional exit (no reference counterpart): a restart after a
    # converged recurrence whose true residual is above STALL_FACTOR times the true
    # residual at the previous such restart ends the solve as "stagnated"
    stop_on_stagnation: bool = False
    # the HIP backend's other optional exit: at most this many restarts after a
    # converged recurrence whose true residual misses the tolerance (None: no limit)
    max_refinements: int | None = None
    # the HIP backend's option to never restart on drift (checks then restart only
    # after a converged recurrence, or at every check with the periodic strategy)
    drift_restarts: bool = True


STALL_FACTOR = 0.5


def pcg_reference(matvec, rhs, apply=None, cfg: Config = Config()) -> Result:
    """solvers.py:167-256."""
    apply = apply or (lambda v: v)
    x = np.zeros_like(rhs)
    r = rhs.copy()
    rhs_norm = float(np.linalg.norm(rhs))
    thr = max(cfg.atol, cfg.rtol * rhs_norm)
    if not math.isfinite(rhs_norm):
        return Result(x, 0, rhs_norm, False, "nonfinite", rhs_norm=rhs_norm)
    if rhs_norm <= thr:
        return Result(x, 0, rhs_norm, True, "residual_tolerance", rhs_norm=rhs_norm)
    z = apply(r)
    rz = float(r @ z)
    d = z.copy()
    it = restarts = checks = 0
    reason = "max_iterations"
    for it in range(1, cfg.max_iterations + 1):
        ad = matvec(d)
        c = float(d @ ad)
        if not (math.isfinite(c) and math.isfinite(rz)):
            reason = "nonfinite"
            break
        if c <= 0 or rz <= 0:
            reason = "nonpositive_curvature"
            break
        a = rz / c
        x = x + a * d
        r = r - a * ad
        rec = float(np.linalg.norm(r))
        periodic = cfg.residual_check_interval > 0 and it % cfg.residual_check_interval == 0
        if rec <= thr or periodic:
            tr = rhs - matvec(x)
            checks += 1
            tn = float(np.linalg.norm(tr))
            if not math.isfinite(tn):
                return Result(x, it, tn, False, "nonfinite", restarts, checks, rhs_norm)
            if tn <= thr:
                return Result(x, it, tn, True, "residual_tolerance", restarts, checks, rhs_norm)
            drift = float(np.linalg.norm(tr - r))
            if rec <= thr or cfg.restart_strategy == "periodic" or drift > cfg.residual_drift_tolerance * max(tn, thr):
                r = tr
                z = apply(r)
                rz = float(r @ z)
                d = z.copy()
                restarts += 1
                continue
        z = apply(r)
        rz_new = float(r @ z)
        d = z + (rz_new / rz) * d
        rz = rz_new
    tn = float(np.linalg.norm(rhs - matvec(x)))
    checks += 1
    conv = math.isfinite(tn) and tn <= thr
    if conv:
        reason = "residual_tolerance"
    elif not math.isfinite(tn):
        reason = "nonfinite"
    return Result(x, it, tn, conv, reason, restarts, checks, rhs_norm)


def pcg_batched(matvec, rhs, apply=None, cfg: Config = Config()) -> Result:
    """warp_pcg.py:110-214 (device status word and batched replays)."""
    apply = apply or (lambda v: v)
    x = np.zeros_like(rhs)
    r = rhs.copy()
    rhs_norm = float(np.linalg.norm(rhs))
    thr = max(cfg.atol, cfg.rtol * rhs_norm)
    if not math.isfinite(rhs_norm) or rhs_norm <= thr:
        fin = math.isfinite(rhs_norm)
        return Result(x, 0, rhs_norm, fin, "residual_tolerance" if fin else "nonfinite", rhs_norm=rhs_norm)
    p = apply(r).copy()
    state = [float(r @ p), thr * thr]
    control = [0, 0, 0]  # after warm-up/capture the reference zeroes control

    def iteration():
        nonlocal x, r, p
        if control[1] != 0 or control[0] >= control[2]:
            frozen = True
        else:
            frozen = False
        ap = matvec(p)
        c = float(p @ ap)
        rz = state[0]
        ok = math.isfinite(c) and math.isfinite(rz) and c > 0 and rz > 0
        if not frozen and ok:
            a = rz / c
            x = x + a * p
            r = r - a * ap
        z = apply(r)
        new_rz = float(r @ z)
        norm2 = float(r @ r)
        if not frozen and ok and math.isfinite(new_rz):
            p = z + (new_rz / rz) * p
        if frozen:
            return
        control[0] += 1
        if not math.isfinite(c) or not math.isfinite(rz) or not math.isfinite(norm2):
            control[1] = 2
        elif c <= 0 or rz <= 0:
            control[1] = 3
        elif norm2 <= state[1]:
            control[1] = 1
        elif not math.isfinite(new_rz):
            control[1] = 2
        state[0] = new_rz

    it = checks = restarts = 0
    reason = "max_iterations"
    tn = math.nan
    converged_true = math.inf
    refinements = 0
    while it < cfg.max_iterations:
        iv = cfg.residual_check_interval
        stop_at = min(cfg.max_iterations, ((it // iv + 1) * iv) if iv else cfg.max_iterations)
        control[1] = 0
        control[2] = stop_at
        for _ in range(min(cfg.graph_batch_iterations, stop_at - it)):
            iteration()
        it, status = control[0], control[1]
        if status == 0 and it < stop_at:
            continue
        tr = rhs - matvec(x)
        tn = float(np.linalg.norm(tr))
        checks += 1
        if not math.isfinite(tn):
            reason = "nonfinite"
            break
        if tn <= thr:
            return Result(x, it, tn, True, "residual_tolerance", restarts, checks, rhs_norm)
        if status in (2, 3):
            reason = "nonfinite" if status == 2 else "nonpositive_curvature"
            break
        if it >= cfg.max_iterations:
            break
        if status == 1:
            if cfg.stop_on_stagnation and tn > STALL_FACTOR * converged_true:
                reason = "stagnated"
                break
            if cfg.max_refinements is not None and refinements >= cfg.max_refinements:
                reason = "stagnated"
                break
            converged_true = tn
            refinements += 1
        drift = float(np.linalg.norm(tr - r))
        drifted = cfg.drift_restarts and drift > cfg.residual_drift_tolerance * max(tn, thr)
        if status == 1 or cfg.restart_strategy == "periodic" or drifted:
            r = tr
            p = apply(r).copy()
            state[0] = float(r @ p)
            restarts += 1
    return Result(x, it, tn, False, reason, restarts, checks, rhs_norm)


def block_jacobi(diagonal: list, layout: list, damping: float):
    """(D_t + damping I)^-1 per block via Cholesky (solvers.py:259-285); raises LinAlgError."""
    invs = []
    for (c, d, off), D in zip(layout, diagonal):
        A = D + damping * np.eye(d)
        L = np.linalg.cholesky(A)
        Li = np.linalg.inv(L)
        invs.append(np.einsum("nki,nkj->nij", Li, Li))

    def apply(v):
        out = np.empty_like(v)
        for (c, d, off), inv in zip(layout, invs):
            seg = v[off : off + c * d].reshape(c, d)
            out[off : off + c * d] = np.einsum("nij,nj->ni", inv, seg).reshape(-1)
        return out

    return apply
