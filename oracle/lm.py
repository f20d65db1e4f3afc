"""float64 Levenberg-Marquardt restatement (CPU oracle) of torchls solvers.solve.

Follows solvers.py:288-696 for the conjugate-gradient and Schur paths
(identity or diagonal damping, optional Jacobi column scaling, no bounds):
initial cost, per-iteration
gradient and termination tests, block-Jacobi PCG (warp_pcg control flow by
default), trial evaluation through expanded differences, gain ratio, and the
fixed-factor / adaptive lambda updates.  ``problem`` provides layout,
linearize, cost, evaluate_trial and retract (see oracle.ba.BAProblem).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .pcg import Config, block_jacobi, pcg_batched, pcg_reference
from .schur import SchurSystem


@dataclass
class TrustRegion:
    lambda_initial: float = 1e-3
    lambda_min: float = 1e-12
    lambda_max: float = 1e12
    lambda_factor: float = 4.0
    step_quality_min: float = 1e-4
    max_trials: int = 20
    strategy: str = "fixed_factor"
    damping_mode: str = "identity"
    diagonal_min: float = 1e-6
    diagonal_max: float = 1e32


@dataclass
class Termination:
    max_iterations: int = 50
    cost_tolerance: float = 1e-10
    gradient_tolerance: float = 1e-8
    parameter_tolerance: float = 1e-10


@dataclass
class Summary:
    initial_cost: float
    final_cost: float
    iterations: int = 0
    accepted_steps: int = 0
    rejected_steps: int = 0
    linear_iterations: int = 0
    termination_reason: str = "max_iterations"
    gradient_norm: float = math.inf
    cost_history: list = field(default_factory=list)
    lambda_history: list = field(default_factory=list)
    linear_iterations_history: list = field(default_factory=list)


def trial_decision(obj, red, pred, lam, rej, tr, usable=True):
    """Gain ratio, acceptance and damping update of one trial (solvers.py:607-617,
    633-655, 683-693): (quality, accepted, damping, rejection factor, stop), where
    stop means the trial loop ends on a rejection (no trust region, or lambda_max
    reached).  usable=False: a trial without a usable step, rejected."""
    q = (red / pred if pred > 0 else -math.inf) if usable else math.nan
    accepted = usable and math.isfinite(obj) and (tr is None or (red > 0 and q >= tr.step_quality_min))
    if accepted:
        if tr:
            if tr.strategy == "adaptive":
                mult = max(1 / 3, 1 - (2 * min(q, 1.0) - 1) ** 3)
                lam = max(tr.lambda_min, min(tr.lambda_max, lam * mult))
                rej = 2.0
            elif q > 0.75:
                lam = max(tr.lambda_min, lam / tr.lambda_factor)
        return q, True, lam, rej, False
    if tr is None or lam >= tr.lambda_max:
        return q, False, lam, rej, True
    if tr.strategy == "adaptive":
        lam, rej = min(tr.lambda_max, lam * rej), min(rej * 2, tr.lambda_max / tr.lambda_min)
    else:
        lam = min(tr.lambda_max, lam * tr.lambda_factor)
    return q, False, lam, rej, False


def _solve_scale(diag, column_scale, jacobi_scaling, tr):
    """(column_scale, solve_scale) of one iteration (solvers.py:430-460): the Jacobi
    column scale 1 / (1 + sqrt(diag(J^T J))) fixed at the first iteration, and with
    diagonal damping the inverse square root of the clamped (scaled) diagonal."""
    if jacobi_scaling and column_scale is None:
        column_scale = np.concatenate(
            [(1.0 / (1.0 + np.sqrt(np.maximum(np.diagonal(b, axis1=1, axis2=2), 0.0)))).reshape(-1) for b in diag])
    solve_scale = column_scale
    if tr is not None and tr.damping_mode == "diagonal":
        nd = np.concatenate([np.diagonal(b, axis1=1, axis2=2).reshape(-1) for b in diag])
        if column_scale is not None:
            nd = nd * column_scale**2
        inv = 1.0 / np.sqrt(np.clip(nd, tr.diagonal_min, tr.diagonal_max))
        solve_scale = inv if column_scale is None else column_scale * inv
    return column_scale, solve_scale


def _scale_blocks(diag, layout, scale):
    """D_t -> S_t D_t S_t per variable block (solvers.py:463-472)."""
    out = []
    for (c, d, off), b in zip(layout, diag):
        s = scale[off : off + c * d].reshape(c, d)
        out.append(b * s[:, :, None] * s[:, None, :])
    return out


def solve(problem, vals, *, linear_solver="conjugate_gradient", cg: Config = Config(),
          tr: TrustRegion | None = TrustRegion(), term: Termination = Termination(), eliminate: int | None = None,
          pcg_flavour: str = "batched", preconditioner: bool = True, jacobi_scaling: bool = False):
    pcg_fn = pcg_batched if pcg_flavour == "batched" else pcg_reference
    c0 = problem.cost(vals)
    s = Summary(c0, c0, cost_history=[c0])
    lam = tr.lambda_initial if tr else 0.0
    rej = 2.0
    column_scale = None
    lin = problem.linearize(vals)
    for it in range(term.max_iterations):
        s.iterations = it + 1
        r, J = lin["residual"], lin["J"]
        if not np.all(np.isfinite(r)) or not all(np.all(np.isfinite(b)) for g in J.groups for b in g.blocks):
            s.termination_reason = "nonfinite_linearization"
            break
        g = J.rmatvec(r)
        s.gradient_norm = float(np.abs(g).max())
        if s.gradient_norm <= term.gradient_tolerance:
            s.termination_reason = "gradient_tolerance"
            break
        J0, g0 = J, g  # the model decrease uses the unscaled system (solvers.py:608-611)
        diag = J.diagonal_blocks()
        column_scale, solve_scale = _solve_scale(diag, column_scale, jacobi_scaling, tr)
        if solve_scale is not None:
            J = J.scale_columns(solve_scale)
            g = g * solve_scale
            diag = _scale_blocks(diag, J.layout, solve_scale)
        old_cost = lin["objective"] if lin["objective"] is not None else 0.5 * float(r @ r)
        accepted = False
        for _ in range(tr.max_trials if tr else 1):
            s.lambda_history.append(lam)
            try:
                if linear_solver == "conjugate_gradient":
                    apply = block_jacobi(diag, J.layout, lam) if preconditioner else None
                    res = pcg_fn(lambda v: J.normal_matvec(v, lam), -g, apply, cg)
                    delta, its = res.solution, res.iterations
                    usable = res.reason in ("residual_tolerance", "max_iterations", "stagnated") and np.all(
                        np.isfinite(delta))
                elif linear_solver in ("schur_cg", "schur_dense"):
                    sys_ = SchurSystem(J, diag, eliminate, lam)
                    rhs = sys_.reduced_rhs(-g)
                    if linear_solver == "schur_dense":
                        S = sys_.dense_matrix()
                        L = np.linalg.cholesky(S)
                        xr = np.linalg.solve(L.T, np.linalg.solve(L, rhs))
                        its = 1
                        usable = bool(np.all(np.isfinite(xr)))
                    else:
                        blocks = sys_.preconditioner_blocks()
                        apply = None
                        if preconditioner:
                            invs = [np.linalg.inv(b) for b in blocks]
                            for b in blocks:
                                np.linalg.cholesky(b)

                            def apply(v, invs=invs):
                                out, k = np.empty_like(v), 0
                                for (c, d, _), inv in zip(sys_.rlayout, invs):
                                    out[k : k + c * d] = np.einsum("nij,nj->ni", inv, v[k : k + c * d].reshape(c, d)).ravel()
                                    k += c * d
                                return out
                        res = pcg_fn(sys_.matvec, rhs, apply, cg)
                        xr, its = res.solution, res.iterations
                        usable = res.reason in ("residual_tolerance", "max_iterations", "stagnated") and np.all(
                            np.isfinite(xr))
                    delta = sys_.backsubstitute(xr)
                elif linear_solver == "dense":
                    A = J.to_dense()
                    delta = np.linalg.solve(A.T @ A + lam * np.eye(J.n), -g)
                    its, usable = 1, True
                else:
                    raise ValueError(linear_solver)
                s.linear_iterations += its
                s.linear_iterations_history.append(its)
            except np.linalg.LinAlgError:
                usable = False
            if usable:
                if solve_scale is not None:
                    delta = solve_scale * delta
                prop = problem.retract(vals, delta)
                obj, red = problem.evaluate_trial(prop, lin)
                pred = float(-(g0 @ delta) - 0.5 * J0.squared_matvec_norm(delta))
                q, accepted, lam_next, rej_next, stop = trial_decision(obj, red, pred, lam, rej, tr)
                if accepted:
                    vals = prop
                    s.accepted_steps += 1
                    s.final_cost = obj
                    s.cost_history.append(obj)
                    lam, rej = lam_next, rej_next
                    lin = problem.linearize(vals)
                    s.gradient_norm = float(np.abs(lin["J"].rmatvec(lin["residual"])).max())
                    if s.gradient_norm <= term.gradient_tolerance:
                        s.termination_reason = "gradient_tolerance"
                    elif abs(red) <= term.cost_tolerance * max(old_cost, np.finfo(np.float64).tiny):
                        s.termination_reason = "cost_tolerance"
                    elif float(np.linalg.norm(delta)) <= term.parameter_tolerance:
                        s.termination_reason = "parameter_tolerance"
                    break
            s.rejected_steps += 1
            if not usable:
                q, accepted, lam_next, rej_next, stop = trial_decision(math.nan, 0.0, 0.0, lam, rej, tr, usable=False)
            if stop:
                break
            lam, rej = lam_next, rej_next
        if not accepted:
            s.termination_reason = "no_acceptable_step"
            break
        if s.termination_reason != "max_iterations":
            break
    return vals, s
