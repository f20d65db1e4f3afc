"""Solver configuration, PCG dispatch, block-Jacobi and the LM / GN loop.

Same interface and control flow as torchls.solvers (solvers.py:19-696) with the
``"hip"`` backend: every operation of the block-sparse normal-equation path --
diagonal blocks, damped block inversion, normal matvec, PCG, the Schur
reduce/matvec/preconditioner/back-substitution, the dense reduced Cholesky, and
the float64 gradient, objective and predicted-reduction reductions, the gain
ratio, acceptance and damping update of each trial -- runs in the HIP kernels
of this package.  ``matvec_dtype=torch.float32`` on float64
parameters selects the mixed-precision mode (float32 blocks, vectors and
matvec; float64 outer residual, gradient, objective, dots and inversions).
"""

from __future__ import annotations

import math
import time
from contextlib import nullcontext
from dataclasses import dataclass, field
from functools import partial
from types import SimpleNamespace

import torch
from torch import Tensor

from . import _lib
from .hip_diagonal import HipBlockJacobi, HipDiagonalPlan, LinAlgError
from .hip_normal import HipNormalPlan
from .hip_ops import LinearizationKernels, lm_decide


@dataclass(frozen=True)
class ConjugateGradientConfig:
    rtol: float = 1e-6
    atol: float = 1e-10
    max_iterations: int = 200
    preconditioner: str | None = "block_jacobi"
    preconditioner_groups: tuple = ()
    residual_check_interval: int = 50
    restart_strategy: str = "adaptive"
    residual_drift_tolerance: float = 0.1
    backend: str = "hip"
    normal_matvec: str = "hip"
    normal_max_packed_bytes: int = 256 * 1024**2
    graph_batch_iterations: int = 16
    cache_retained_matrix: bool = False
    # Storage/arithmetic dtype of the packed blocks and PCG vectors; None = the
    # parameters' dtype.  torch.float32 with float64 parameters = mixed precision.
    matvec_dtype: torch.dtype | None = None
    # Not in the reference (warp_pcg keeps restarting until max_iterations): when set,
    # a restart after a converged recurrence that fails to halve the true residual of
    # the previous such restart ends the solve with reason "stagnated" (a usable
    # step, like "max_iterations").  This is where a float32 matvec under float64
    # PCG vectors reaches its rounding floor.
    stop_on_stagnation: bool = False
    # Not in the reference either: when set, at most this many restarts follow a
    # converged recurrence whose true residual misses the tolerance; the next such
    # check ends the solve as "stagnated".  0 suits a float32 operator, whose true
    # residual floor usually lies above rtol: the outer float64 iteration corrects
    # the step instead of repeated refinement inside the solve.
    max_refinements: int | None = None
    # Not in the reference either: when False, a residual check restarts only after a
    # converged recurrence (or at every check with the periodic strategy), never on
    # drift.  With float32 PCG vectors the computed true residual carries a rounding
    # floor that the drift test cannot tell from drift of the recurrence: on an
    # ill-conditioned system (a long pose-graph chain) it then restarts at every check
    # and discards the Krylov space, while the float64 outer iteration corrects the
    # step anyway.  Float64 solves rarely drift, so False usually changes nothing there.
    drift_restarts: bool = True

    def __post_init__(self):
        if (isinstance(self.normal_max_packed_bytes, bool) or not isinstance(self.normal_max_packed_bytes, int)
                or self.normal_max_packed_bytes < 0):
            raise ValueError("normal_max_packed_bytes must be a nonnegative integer")
        if (isinstance(self.graph_batch_iterations, bool) or not isinstance(self.graph_batch_iterations, int)
                or self.graph_batch_iterations < 1):
            raise ValueError("graph_batch_iterations must be a positive integer")
        if self.normal_matvec != "hip":
            raise ValueError("normal_matvec must be hip (this package implements the HIP backend)")
        if self.backend != "hip":
            raise ValueError("CG backend must be hip (this package implements the HIP backend)")
        if self.rtol < 0 or self.atol < 0 or self.max_iterations < 1:
            raise ValueError("Invalid CG tolerance or iteration limit")
        if self.preconditioner not in ("block_jacobi", None):
            raise ValueError("preconditioner must be block_jacobi or None")
        from .preconditioning import normalize_groups

        object.__setattr__(self, "preconditioner_groups", normalize_groups(self.preconditioner_groups))
        if self.preconditioner_groups and self.preconditioner is None:
            raise ValueError("Joint groups require block_jacobi preconditioning")
        if (isinstance(self.residual_check_interval, bool) or not isinstance(self.residual_check_interval, int)
                or self.residual_check_interval < 0):
            raise ValueError("residual_check_interval must be a nonnegative integer")
        if self.restart_strategy not in ("adaptive", "periodic"):
            raise ValueError("restart_strategy must be adaptive or periodic")
        if not math.isfinite(self.residual_drift_tolerance) or self.residual_drift_tolerance <= 0:
            raise ValueError("residual_drift_tolerance must be finite and positive")
        if self.matvec_dtype not in (None, torch.float32, torch.float64):
            raise ValueError("matvec_dtype must be None, torch.float32 or torch.float64")
        if not isinstance(self.stop_on_stagnation, bool):
            raise ValueError("stop_on_stagnation must be a bool")
        if not isinstance(self.drift_restarts, bool):
            raise ValueError("drift_restarts must be a bool")
        if self.max_refinements is not None and (
                isinstance(self.max_refinements, bool) or not isinstance(self.max_refinements, int)
                or self.max_refinements < 0):
            raise ValueError("max_refinements must be None or a nonnegative integer")


@dataclass(frozen=True)
class TrustRegionConfig:
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

    def __post_init__(self):
        if not (0 < self.lambda_min <= self.lambda_initial <= self.lambda_max):
            raise ValueError("Require 0 < lambda_min <= lambda_initial <= lambda_max")
        if self.lambda_factor <= 1 or self.max_trials < 1 or not (0 <= self.step_quality_min < 1):
            raise ValueError("Invalid trust region settings")
        if self.strategy not in ("fixed_factor", "adaptive"):
            raise ValueError("Trust region strategy must be fixed_factor or adaptive")
        if self.damping_mode not in ("identity", "diagonal"):
            raise ValueError("damping_mode must be identity or diagonal")
        if not (0 < self.diagonal_min <= self.diagonal_max < math.inf):
            raise ValueError("Require finite positive diagonal damping bounds")


@dataclass(frozen=True)
class TerminationConfig:
    max_iterations: int = 50
    cost_tolerance: float = 1e-10
    gradient_tolerance: float = 1e-8
    parameter_tolerance: float = 1e-10
    target_cost: float | None = None
    max_time_seconds: float | None = None

    def __post_init__(self):
        if self.max_iterations < 1 or min(self.cost_tolerance, self.gradient_tolerance, self.parameter_tolerance) < 0:
            raise ValueError("Invalid termination settings")
        if self.target_cost is not None and (not math.isfinite(self.target_cost) or self.target_cost < 0):
            raise ValueError("target_cost must be finite and nonnegative")
        if self.max_time_seconds is not None and (
            not math.isfinite(self.max_time_seconds) or self.max_time_seconds <= 0
        ):
            raise ValueError("max_time_seconds must be finite and positive")


_DEFAULT_CG = ConjugateGradientConfig()
_DEFAULT_TRUST_REGION = TrustRegionConfig()
_DEFAULT_TERMINATION = TerminationConfig()


@dataclass(frozen=True)
class PCGResult:
    solution: Tensor
    iterations: int
    residual_norm: float
    converged: bool
    reason: str
    restarts: int = 0
    residual_checks: int = 0
    rhs_norm: float = 0.0


@dataclass
class SolveSummary:
    """Histories include the initial point and each accepted update (solvers.py:142-164)."""

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
    linear_converged_history: list = field(default_factory=list)
    trial_history: list = field(default_factory=list)
    linear_solver: str = "conjugate_gradient"
    jacobi_scaling: bool = False
    reduced_dim: int | None = None
    linear_restarts: int = 0
    linear_residual_checks: int = 0
    linear_relative_residual_history: list = field(default_factory=list)


def pcg(matvec, rhs: Tensor, *, preconditioner=None, config: ConjugateGradientConfig = _DEFAULT_CG,
        first_batch: int | None = None) -> PCGResult:
    """Preconditioned CG with true-residual checks (solvers.py:167-256), on the HIP backend.

    ``first_batch``: iterations run before the first host read (default
    ``config.graph_batch_iterations``); the LM loop passes the previous solve's count.
    """
    if rhs.ndim != 1 or rhs.numel() == 0:
        raise ValueError("PCG expects a nonempty vector")
    from .hip_pcg import hip_pcg

    return hip_pcg(matvec, rhs, preconditioner=preconditioner, config=config, first_batch=first_batch)


def block_jacobi(jacobian, damping: float, *, diagonal=None, groups=(), graph_safe=False):
    """float64 (D_t + damping I)^-1 per variable block (solvers.py:259-285).

    Raises ``LinAlgError`` when a damped block is not positive definite."""
    if diagonal is None:
        diagonal = HipDiagonalPlan(jacobian).diagonal_blocks(jacobian)
    if groups:
        from .preconditioning import BlockJacobiPlan

        plan = BlockJacobiPlan(jacobian.layout, groups, jacobian.groups[0].blocks[0].device)
        return plan.factor(plan.normal_blocks(jacobian, diagonal), damping)
    return HipBlockJacobi(jacobian.layout, tuple(d.to(torch.float64) for d in diagonal), damping)


@torch.no_grad()
def solve(
    problem,
    initial_vals,
    *,
    linear_solver: str = "conjugate_gradient",
    conjugate_gradient: ConjugateGradientConfig = _DEFAULT_CG,
    trust_region: TrustRegionConfig | None = _DEFAULT_TRUST_REGION,
    termination: TerminationConfig = _DEFAULT_TERMINATION,
    return_summary: bool = False,
    jacobi_scaling: bool = False,
    eliminate=None,
    profile: bool = False,
    record_trials: bool = False,
    bounds=None,
    normal_blocks_backend: str = "auto",
):
    """Levenberg-Marquardt (Gauss-Newton when trust_region=None) on the HIP path.

    ``problem`` is an ``Analysis`` (from ``Problem``); ``initial_vals`` maps
    parameter names to tensors.  Returns the solved dict (and the summary).
    """
    if linear_solver not in ("conjugate_gradient", "dense_qr", "schur_dense", "schur_cg"):
        raise ValueError("Unknown linear solver")
    if normal_blocks_backend not in ("auto", "hip"):
        raise ValueError("normal_blocks_backend must be auto or hip")
    if bounds is not None:
        raise ValueError("bounds are not supported by the HIP backend")
    is_schur = linear_solver.startswith("schur_")
    if eliminate is not None and not is_schur:
        raise ValueError("eliminate requires a Schur linear solver")
    if conjugate_gradient.preconditioner_groups and linear_solver not in ("conjugate_gradient", "schur_cg"):
        raise ValueError("Joint preconditioner groups require a CG linear solver")
    if is_schur and eliminate is None:
        raise ValueError("Schur solvers need eliminate=<parameter group name>")
    start = time.perf_counter()

    def region(name):
        return torch.profiler.record_function("torchls/" + name) if profile else nullcontext()

    def timed_out():
        return termination.max_time_seconds is not None and time.perf_counter() - start >= termination.max_time_seconds

    vals = {k: v.detach() for k, v in initial_vals.items()}
    first = next(iter(vals.values()))
    device, dtype = first.device, first.dtype
    if device.type != "cuda":
        raise ValueError("the HIP backend solves on a GPU device; move the parameters to cuda")
    mv_dtype = conjugate_gradient.matvec_dtype or dtype
    # mixed precision: the Jacobian is stored in the matvec dtype as it is linearized
    # (float32 blocks, float64 residuals); the gradient, diagonal blocks and model
    # decrease are still summed in float64 from those blocks
    block_dtype = mv_dtype if mv_dtype != dtype else None
    initial_cost = float(problem.cost(vals))
    if not math.isfinite(initial_cost):
        raise ValueError("Initial residuals are nonfinite")
    summary = SolveSummary(initial_cost, initial_cost, cost_history=[initial_cost], linear_solver=linear_solver,
                           jacobi_scaling=jacobi_scaling)
    damping = trust_region.lambda_initial if trust_region else 0.0
    # the gain ratio, acceptance and damping update of every trial run on the device
    # (tlh_lm_decide), which keeps [damping, rejection factor]; `damping` mirrors it
    lm_trust = _lib.LmTrust(
        1 if trust_region else 0, 1 if trust_region and trust_region.strategy == "adaptive" else 0,
        trust_region.lambda_min if trust_region else 0.0, trust_region.lambda_max if trust_region else 0.0,
        trust_region.lambda_factor if trust_region else 0.0, trust_region.step_quality_min if trust_region else 0.0)
    lm_state = torch.tensor([damping, 2.0], dtype=torch.float64, device=device)
    # Symbolic plans depend only on the residual structure, which an Analysis
    # fixes: they are built on the first solve and reused by later solves.
    plan_cache = getattr(problem, "plans", None)

    def cached(key, make):
        if plan_cache is None:
            return make()
        key = key + (str(device), dtype, mv_dtype)
        if key not in plan_cache:
            plan_cache[key] = make()
        return plan_cache[key]

    schur_plan = None
    if is_schur:
        from .schur import HipSchurPlan

        with region("analyze_schur"):
            groups = conjugate_gradient.preconditioner_groups
            schur_plan = cached(("schur", eliminate, groups),
                                lambda: HipSchurPlan(problem, eliminate, device, dtype=mv_dtype,
                                                     preconditioner_groups=groups))
        if linear_solver == "schur_dense" and schur_plan.dim > 8192:
            raise ValueError("Dense Schur is limited to 8192 retained dimensions; use schur_cg")
        summary.reduced_dim = schur_plan.dim
    normal_plan = diagonal_plan = kernels = None
    column_scale = None
    joint_plan = None
    if conjugate_gradient.preconditioner_groups and linear_solver == "conjugate_gradient":
        from .preconditioning import BlockJacobiPlan

        joint_plan = cached(("joint", conjugate_gradient.preconditioner_groups),
                            lambda: BlockJacobiPlan(problem.layout, conjugate_gradient.preconditioner_groups, device))

    def examine_device(lin):
        """float64 gradient of a linearization and its (nonfinite count, max |g|,
        0.5 sum rho) as three one-element device tensors, read by the caller with
        other values."""
        nonlocal diagonal_plan, kernels
        jac = lin.jacobian
        if kernels is None:
            diagonal_plan = cached(("diagonal",), lambda: HipDiagonalPlan(jac, dtype=mv_dtype))
            kernels = cached(("kernels",), lambda: LinearizationKernels(jac, diagonal_plan))
        g = kernels.gradient(jac, lin.residual)
        cost = lin.objective if lin.objective is not None else 0.5 * lin.residual.to(torch.float64).square().sum()
        bad = lin.nonfinite if lin.nonfinite is not None else (~torch.isfinite(lin.residual)).sum().double()
        return g, [bad.to(torch.float64).reshape(1), kernels.vector_stats(g, None)[3:], cost.to(torch.float64).reshape(1)]

    def examine(lin):
        """examine_device plus its host read: (gradient, finite, max |g|, cost)."""
        g, stats = examine_device(lin)
        host = torch.cat(stats).cpu()
        return g, bool(host[0] == 0), float(host[1]), float(host[2])

    def prepare(jacobian, gradient):
        """The damping-independent part of a linear system: diagonal blocks, column scale,
        scaled gradient and the packed Schur or normal operator of one linearization."""
        nonlocal column_scale, normal_plan
        diagonal_damping = trust_region is not None and trust_region.damping_mode == "diagonal"
        needs_diagonal = is_schur or (linear_solver == "conjugate_gradient"
                                      and conjugate_gradient.preconditioner is not None)
        diagonal = None
        if needs_diagonal or diagonal_damping or (jacobi_scaling and column_scale is None):
            packed_unscaled = diagonal_plan._pack(jacobian)
            diagonal = diagonal_plan.diagonal_blocks_f64(packed_unscaled, jacobian)
        if jacobi_scaling and column_scale is None:
            column_scale = torch.cat(
                [(1 / (1 + b.diagonal(dim1=-2, dim2=-1).clamp_min(0).sqrt())).reshape(-1) for b in diagonal]
            )
        solve_scale = column_scale
        if diagonal_damping:
            nd = torch.cat([b.diagonal(dim1=-2, dim2=-1).reshape(-1) for b in diagonal])
            if column_scale is not None:
                nd = nd * column_scale.square()
            inv_diag = nd.clamp(trust_region.diagonal_min, trust_region.diagonal_max).rsqrt()
            solve_scale = inv_diag if column_scale is None else column_scale * inv_diag
        if solve_scale is not None:
            # J diag(s) is never copied: the pack kernels scale the blocks as they
            # lay them out for the normal / Schur operators (tlh_pack_blocks_scaled)
            solve_scale = solve_scale.to(torch.float64).contiguous()
            linear_gradient = gradient * solve_scale
            if diagonal is not None:
                diagonal = tuple(
                    b * solve_scale[e.offset: e.offset + e.count * e.dim].reshape(e.count, e.dim, 1)
                    * solve_scale[e.offset: e.offset + e.count * e.dim].reshape(e.count, 1, e.dim)
                    for e, b in zip(problem.layout, diagonal)
                )
        else:
            linear_gradient = gradient
        schur_linear = normal_operator = None
        joint_diagonal = (joint_plan.normal_blocks(jacobian, diagonal, col_scale=solve_scale)
                          if joint_plan is not None else None)
        if is_schur:
            schur_linear = schur_plan.linearize(jacobian, diagonal, dense=linear_solver == "schur_dense",
                                                col_scale=solve_scale)
        elif linear_solver == "conjugate_gradient":
            if normal_plan is None:
                normal_plan = cached(("normal",), lambda: HipNormalPlan(jacobian, dtype=mv_dtype))
            normal_operator = normal_plan.bind(jacobian, col_scale=solve_scale)
        return SimpleNamespace(jacobian=jacobian, diagonal=diagonal, solve_scale=solve_scale,
                               linear_gradient=linear_gradient, schur_linear=schur_linear,
                               normal_operator=normal_operator, joint_diagonal=joint_diagonal)

    with region("linearize"):
        lin = problem.linearize(vals, block_dtype=block_dtype)
    head = None
    prepared = None  # the prepared system of `lin`, when issued ahead (speculatively, before the trial read)
    pcg_hint = None  # iterations of the previous PCG solve: the next one's first batch
    for iteration in range(termination.max_iterations):
        if timed_out():
            summary.termination_reason = "time_limit"
            break
        summary.iterations = iteration + 1
        residual, jacobian = lin.residual, lin.jacobian
        if head is None:
            head = examine(lin)
        gradient, finite, summary.gradient_norm, old_cost = head
        head = None
        if not finite:
            summary.termination_reason = "nonfinite_linearization"
            break
        if termination.target_cost is not None and old_cost <= termination.target_cost:
            summary.termination_reason = "target_cost"
            break
        if summary.gradient_norm <= termination.gradient_tolerance:
            summary.termination_reason = "gradient_tolerance"
            break
        original_jacobian = jacobian
        if prepared is None or prepared.jacobian is not jacobian:
            with region("prepare_linear_system"):
                prepared = prepare(jacobian, gradient)
        diagonal, solve_scale, linear_gradient = prepared.diagonal, prepared.solve_scale, prepared.linear_gradient
        schur_linear, normal_operator, joint_diagonal = (prepared.schur_linear, prepared.normal_operator,
                                                         prepared.joint_diagonal)
        prepared = None
        accepted = False
        trial_limit = trust_region.max_trials if trust_region else 1
        for trial in range(trial_limit):
            if timed_out():
                summary.termination_reason = "time_limit"
                break
            summary.lambda_history.append(damping)
            # max |step| of the iterative solves (float64, one reduction), read with the trial:
            # finite exactly when every entry is
            finite_step = None
            try:
                if linear_solver == "dense_qr":
                    dense = jacobian.to_dense().to(dtype)
                    if solve_scale is not None:
                        dense = dense * solve_scale.to(dense.dtype)
                    eye = torch.eye(problem.tangent_dim, dtype=dtype, device=device)
                    aug = torch.cat((dense, math.sqrt(damping) * eye))
                    rhs = torch.cat((-residual, gradient.new_zeros(problem.tangent_dim).to(residual.dtype)))
                    delta = torch.linalg.lstsq(aug, rhs[:, None]).solution[:, 0]
                    usable = bool(torch.isfinite(delta).all())
                elif is_schur:
                    with region("prepare_schur"):
                        # with a preconditioner, the PCG's first read also checks the damped
                        # blocks' positive definiteness (LinAlgError before any iteration)
                        system = schur_linear.damped(damping, defer_check=linear_solver == "schur_cg" and
                                                     conjugate_gradient.preconditioner is not None)
                        rhs = system.reduced_rhs(linear_gradient, scale=-1.0)
                    with region("linear_solve"):
                        if linear_solver == "schur_dense":
                            retained = system.dense_solve(rhs)
                            iterations, converged = 1, True
                            usable = bool(torch.isfinite(retained).all())
                            summary.linear_relative_residual_history.append(None)
                        else:
                            result = pcg(system.matvec, rhs,
                                         preconditioner=system.preconditioner() if conjugate_gradient.preconditioner
                                         else None, config=conjugate_gradient, first_batch=pcg_hint)
                            pcg_hint = max(1, result.iterations)
                            retained, iterations, converged = result.solution, result.iterations, result.converged
                            # the step's finiteness is read with the trial (one host read per trial)
                            usable = result.reason in ("residual_tolerance", "max_iterations", "stagnated")
                            finite_step = kernels.vector_stats(None, retained)[2:3]  # max |retained|
                            summary.linear_restarts += result.restarts
                            summary.linear_residual_checks += result.residual_checks
                            summary.linear_relative_residual_history.append(
                                result.residual_norm / result.rhs_norm if result.rhs_norm else result.residual_norm)
                        delta = system.backsubstitute(retained)
                    summary.linear_iterations += iterations
                    summary.linear_iterations_history.append(iterations)
                    summary.linear_converged_history.append(converged)
                else:
                    if conjugate_gradient.preconditioner is None:
                        precondition = None
                    elif joint_plan is not None:
                        precondition = joint_plan.factor(joint_diagonal, damping, defer_check=True)
                    else:
                        precondition = HipBlockJacobi(problem.layout, diagonal, damping, defer_check=True)
                    with region("linear_solve"):
                        result = pcg(partial(normal_operator.normal_matvec, damping=damping),
                                     (-linear_gradient).to(mv_dtype), preconditioner=precondition,
                                     config=conjugate_gradient, first_batch=pcg_hint)
                        pcg_hint = max(1, result.iterations)
                    summary.linear_iterations += result.iterations
                    summary.linear_iterations_history.append(result.iterations)
                    summary.linear_converged_history.append(result.converged)
                    summary.linear_restarts += result.restarts
                    summary.linear_residual_checks += result.residual_checks
                    summary.linear_relative_residual_history.append(
                        result.residual_norm / result.rhs_norm if result.rhs_norm else result.residual_norm)
                    delta = result.solution
                    usable = result.reason in ("residual_tolerance", "max_iterations", "stagnated")
            except (LinAlgError, torch.linalg.LinAlgError):
                usable = False
            if usable:
                delta = delta.to(torch.float64)
                if solve_scale is not None:
                    delta = solve_scale.to(torch.float64) * delta
                proposed = problem.retract(vals, delta, kernel=True)
                with region("trial_residual"):
                    objective, reduction = problem.evaluate_trial(proposed, lin)
                sq = kernels.squared_matvec_norm(original_jacobian, delta)
                # (g.delta, delta.delta, max |delta|, -) in one launch; the full-system step's
                # finiteness is that of max |delta| (the solution times a finite positive scale)
                step = kernels.vector_stats(gradient, delta)
                if finite_step is None:  # dense solves checked their step before the trial
                    finite_step = step[2:3] if linear_solver == "conjugate_gradient" else step.new_zeros(1)
                # predicted reduction, gain ratio, acceptance and the damping update, on the device
                decision = lm_decide((objective.reshape(1), reduction.reshape(1), sq.reshape(1), step,
                                              finite_step), lm_state, lm_trust)
                parts = [objective.reshape(1), reduction.reshape(1), decision]
                # The next linearization is issued before the trial is read (used if the step
                # is accepted, dropped otherwise): its gradient and cost ride on the same read.
                with region("linearize"):
                    lin_next = problem.linearize(proposed, block_dtype=block_dtype)
                g_next, stats_next = examine_device(lin_next)
                # The first trial of an iteration also issues the next system's damping-
                # independent part (kept if the step is accepted); later trials, which
                # follow a rejection, do not.
                prep_next = None
                if trial == 0 and iteration + 1 < termination.max_iterations and linear_solver != "dense_qr":
                    with region("prepare_linear_system"):
                        prep_next = prepare(lin_next.jacobian, g_next)
                host = torch.cat(parts + stats_next).cpu().tolist()
                proposed_cost, actual = host[:2]
                verdict = host[2:11]
                stats_host = host[11:14]  # the next linearization's (nonfinite count, max |g|, cost)
            else:  # no usable step: a rejection, the damping update still on the device
                verdict = lm_decide(None, lm_state, lm_trust).cpu().tolist()
            predicted, quality, accepted, usable, step_norm, _, stop, damping_next, _ = verdict
            accepted, usable, stop = accepted != 0.0, usable != 0.0, stop != 0.0
            if usable:  # a nonfinite step is rejected without a trial record
                if record_trials:
                    summary.trial_history.append(dict(damping=damping, predicted_reduction=predicted,
                                                      actual_reduction=actual, gain_ratio=quality,
                                                      accepted=accepted))
                if accepted:
                    vals = proposed
                    summary.accepted_steps += 1
                    summary.final_cost = proposed_cost
                    summary.cost_history.append(proposed_cost)
                    damping = damping_next
                    residual = jacobian = original_jacobian = None
                    normal_operator = schur_linear = system = diagonal = gradient = linear_gradient = None
                    joint_diagonal = None
                    lin = lin_next
                    prepared = prep_next
                    head = (g_next, stats_host[0] == 0, stats_host[1], stats_host[2])
                    summary.gradient_norm = head[2]
                    if termination.target_cost is not None and proposed_cost <= termination.target_cost:
                        summary.termination_reason = "target_cost"
                    elif summary.gradient_norm <= termination.gradient_tolerance:
                        summary.termination_reason = "gradient_tolerance"
                    elif abs(actual) <= termination.cost_tolerance * max(old_cost, torch.finfo(dtype).tiny):
                        summary.termination_reason = "cost_tolerance"
                    elif step_norm <= termination.parameter_tolerance:
                        summary.termination_reason = "parameter_tolerance"
                    break
            summary.rejected_steps += 1
            if stop:  # rejected at lambda_max, or without a trust region
                break
            damping = damping_next
        if not accepted:
            if summary.termination_reason != "time_limit":
                summary.termination_reason = "no_acceptable_step"
            break
        if summary.termination_reason != "max_iterations":
            break
    return (vals, summary) if return_summary else vals
