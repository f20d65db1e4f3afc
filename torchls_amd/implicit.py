"""Implicit differentiation of the LM optimum with respect to observations.

``solve_implicit(problem, observations, ...)`` solves like ``problem.solve()``
and returns the solved parameter tensors as differentiable functions of the
``observations`` tensor (a data argument of one reprojection residual batch).
Backward (no reference counterpart; the reference solve is
``@torch.no_grad``, solvers.py:288, 307-308):

1. the cotangents of the solved tensors are pulled back to the tangent space
   through the retraction at delta = 0;
2. one adjoint PCG solve ``(J~^T J~ + lambda I) w = dL/dtheta`` on the solver's
   own HIP operator at the optimum (HIP normal matvec, block Jacobi, HIP PCG);
   ``refine > 0`` continues with up to ``refine`` conjugate-gradient steps on
   the exact Hessian, preconditioned by that Gauss-Newton solve, which removes
   the Gauss-Newton approximation error (large for Huber outliers, whose exact
   curvature along their residual vanishes).  The exact Hessian-vector product
   of the reprojection batches runs in the HIP kernel ``tlh_reprojection_hvp``
   (hyper-dual numbers through the retraction and the projection, summed per
   variable by ``tlh_segment_sum_f64``); other residual batches (priors and
   user residuals) use ``torch.func``;
3. the HIP backward kernel ``tlh_reprojection_obs_vjp`` recomputes each
   observation's Jacobian blocks and applies the residual-vs-observation
   Jacobian transpose: ``dL/dobs_i = (rho' I + 2 rho'' r r^T)(J_i w)``.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from .hip_diagonal import HipBlockJacobi, HipDiagonalPlan
from .hip_normal import HipNormalPlan
from .hip_pcg import hip_pcg
from .losses import HuberLoss
from .residuals import reprojection
from .solvers import ConjugateGradientConfig


def _find_batch(problem, observations):
    for k, (fn, args, opts) in enumerate(problem._calls):
        for pos, a in enumerate(args):
            if a is observations:
                if fn is not reprojection:
                    raise NotImplementedError("implicit gradients are implemented for the reprojection residual")
                return k, pos, opts
    raise ValueError("observations must be a data tensor passed to Problem.add_residual")


def _tangent_gradient(analysis, values, grads):
    """dL/ddelta at delta = 0 for every layout entry (vjp through the retraction)."""
    parts = []
    for e in analysis.layout:
        x = values[e.name]
        g = grads.get(e.name)
        if g is None:
            parts.append(torch.zeros(e.count * e.dim, dtype=torch.float64, device=x.device))
            continue
        zeros = torch.zeros(e.count, e.dim, dtype=x.dtype, device=x.device)

        def one(xi, gi, di, fn=e.var_type.retract_fn):
            return torch.func.vjp(lambda d: fn(xi, d), di)[1](gi)[0]

        parts.append(torch.func.vmap(one)(x, g.to(x.dtype), zeros).reshape(-1).to(torch.float64))
    return torch.cat(parts)


def _objective(analysis, values, delta, batches=None):
    """float64 0.5 * sum rho(||r||^2) at retract(values, delta) (differentiable), over
    ``batches`` (indices; default all)."""
    moved = analysis.retract(values, delta)
    first = next(iter(moved.values()))
    total = torch.zeros((), dtype=torch.float64, device=first.device)
    for k, b in enumerate(analysis.batches):
        if batches is not None and k not in batches:
            continue
        r = b.residuals(moved, analysis.layout, first.device, first.dtype).to(torch.float64)
        s = r.square().sum(-1)
        v = s if b.loss is None else b.loss.value(s)
        w = b.device_state(first.device, analysis.layout)[2]
        if w is not None:
            v = v * w
        total = total + 0.5 * v.sum()
    return total


def _csr(index: torch.Tensor, count: int):
    """(order, ptr) int32 grouping of rows by variable index (stable)."""
    order = torch.argsort(index, stable=True)
    ptr = torch.zeros(count + 1, dtype=torch.int64, device=index.device)
    ptr[1:] = torch.bincount(index, minlength=count).cumsum(0)
    return order.to(torch.int32).contiguous(), ptr.to(torch.int32).contiguous()


class ExactHessian:
    """v -> (H + damping I) v, H the exact Hessian of the float64 objective
    0.5 sum w rho(||r||^2) composed with the retraction, at ``values`` (delta = 0).

    Reprojection batches with float64 4x4 poses run in HIP (``tlh_reprojection_hvp``
    per observation, ``tlh_segment_sum_f64`` per camera and point, fixed order);
    every other batch through ``torch.func`` (forward-over-reverse)."""

    def __init__(self, analysis, values, damping: float = 0.0):
        self.analysis, self.values, self.damping = analysis, values, float(damping)
        first = next(iter(values.values()))
        self.device = first.device
        self.lib = _lib.load()
        self.hip, others = [], []
        for k, b in enumerate(analysis.batches):
            ops = b.hip_operands(values, analysis.layout, first.device, first.dtype)
            if ops is None or first.dtype != torch.float64 or ops[0].shape[1:] != (4, 4):
                others.append(k)
                continue
            cams, pts, obs, delta = ops
            idx, _, w = b.device_state(first.device, analysis.layout)
            tc, tp = b.type_indices
            ec, ep = analysis.layout[tc], analysis.layout[tp]
            self.hip.append(dict(cams=cams, pts=pts, obs=obs, delta=delta, ci=idx[0].contiguous(),
                                 pi=idx[1].contiguous(), w=None if w is None else w.contiguous(), count=b.count,
                                 cam=(ec.offset, ec.count) + _csr(idx[0], ec.count),
                                 pt=(ep.offset, ep.count) + _csr(idx[1], ep.count)))
        self.others = others

    def __call__(self, v: torch.Tensor) -> torch.Tensor:
        v = v.to(torch.float64).contiguous()
        out = self.damping * v
        stream = _lib.stream_handle(self.device)
        for h in self.hip:
            rows = torch.empty((h["count"], 9), dtype=torch.float64, device=self.device)
            _lib.check(self.lib.tlh_reprojection_hvp(
                h["cams"].data_ptr(), h["pts"].data_ptr(), h["obs"].data_ptr(), h["ci"].data_ptr(),
                h["pi"].data_ptr(), v.data_ptr(), h["cam"][0], h["pt"][0], h["delta"],
                None if h["w"] is None else h["w"].data_ptr(), h["count"], rows.data_ptr(), stream),
                "tlh_reprojection_hvp")
            for (off, count, order, ptr), col0, dim in ((h["cam"], 0, 6), (h["pt"], 6, 3)):
                part = torch.empty(count * dim, dtype=torch.float64, device=self.device)
                _lib.check(self.lib.tlh_segment_sum_f64(rows.data_ptr(), 9, col0, dim, order.data_ptr(),
                                                        ptr.data_ptr(), count, part.data_ptr(), stream),
                           "tlh_segment_sum_f64")
                out[off: off + count * dim] += part
        if self.others:
            zero = torch.zeros_like(v)
            batches = set(self.others)
            with torch.enable_grad():
                hv = torch.func.jvp(torch.func.grad(lambda d: _objective(self.analysis, self.values, d, batches)),
                                    (zero,), (v,))[1]
            out = out + hv
        return out


def _gn_solver(analysis, values, damping, config):
    """v -> (J~^T J~ + damping I)^-1 v by HIP PCG (float64 normal matvec, block Jacobi) at values.
    The symbolic plans are cached on the analysis, like the solver's."""
    lin = analysis.linearize(values)
    J = lin.jacobian
    key = ("implicit", str(J.groups[0].blocks[0].device))
    if key not in analysis.plans:
        analysis.plans[key] = (HipNormalPlan(J, dtype=torch.float64), HipDiagonalPlan(J))
    normal_plan, diagonal_plan = analysis.plans[key]
    op = normal_plan.bind(J)
    diag = diagonal_plan.diagonal_blocks(J)
    pre = HipBlockJacobi(J.layout, tuple(d.double() for d in diag), damping)
    matvec = lambda v: op.normal_matvec(v, damping)  # noqa: E731
    return lin, (lambda v: hip_pcg(matvec, v, preconditioner=pre, config=config).solution)


def newton_polish(problem, *, steps: int = 3, iterations: int = 50, rtol: float = 1e-13,
                  config: ConjugateGradientConfig | None = None):
    """Refine the problem's parameters (near an optimum) by Newton steps on the exact
    float64 Hessian: each step solves H delta = -g to relative residual ``rtol`` by
    conjugate gradients with ``ExactHessian`` products, preconditioned by the
    Gauss-Newton HIP PCG solve, then retracts.  Returns (and stores) the refined
    parameters."""
    analysis = problem._analyze()
    cfg = config or ConjugateGradientConfig(rtol=rtol, atol=0.0, max_iterations=2000, residual_check_interval=25)
    vals = {k: v.detach() for k, v in problem.parameters.items()}
    with torch.no_grad():
        for _ in range(steps):
            lin, gn = _gn_solver(analysis, vals, 0.0, cfg)
            g = lin.jacobian.rmatvec(lin.residual).to(torch.float64)
            H = ExactHessian(analysis, vals)
            delta = _exact_adjoint(H, gn, -g, gn(-g), iterations, rtol=rtol)
            vals = analysis.retract(vals, delta, kernel=True)
    problem.parameters.update(vals)
    return dict(vals)


def _exact_adjoint(hvp, precondition, b, w, iterations, rtol=1e-12):
    """Conjugate gradients on H w = b from ``w`` (H = exact Hessian through ``hvp``),
    preconditioned by the Gauss-Newton solve ``precondition``."""
    r = b - hvp(w)
    z = precondition(r)
    p = z.clone()
    rz = torch.dot(r, z)
    bound = rtol * float(torch.linalg.vector_norm(b))
    for _ in range(iterations):
        if float(torch.linalg.vector_norm(r)) <= bound:
            break
        q = hvp(p)
        alpha = rz / torch.dot(p, q)
        w = w + alpha * p
        r = r - alpha * q
        z = precondition(r)
        rz_next = torch.dot(r, z)
        p = z + (rz_next / rz) * p
        rz = rz_next
    return w


class _ImplicitSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, observations, state):
        problem, options = state["problem"], state["options"]
        with torch.no_grad():
            sol, summary = problem.solve(return_summary=True, **options)
            if state["polish"] > 0:
                sol = newton_polish(problem, steps=state["polish"])
        state["summary"] = summary
        state["solution"] = {k: v.detach().clone() for k, v in sol.items()}
        ctx.state = state
        names = list(sol)
        state["names"] = names
        return tuple(sol[k].detach().clone() for k in names)

    @staticmethod
    def backward(ctx, *grads):
        st = ctx.state
        problem = st["problem"]
        analysis = problem._analyze()
        values = st["solution"]
        grad_map = {k: g for k, g in zip(st["names"], grads) if g is not None}
        b = _tangent_gradient(analysis, values, grad_map)
        lam = st["adjoint_damping"]
        _, gn_solve = _gn_solver(analysis, values, lam, st["adjoint_config"])
        w = gn_solve(b)
        if st["refine"] > 0:
            w = _exact_adjoint(ExactHessian(analysis, values, lam), gn_solve, b, w, st["refine"])
        k, pos, opts = st["batch"], st["arg"], st["opts"]
        batch = analysis.batches[k]
        cams = values[analysis.layout[batch.type_indices[0]].name].to(torch.float64).contiguous()
        pts = values[analysis.layout[batch.type_indices[1]].name].to(torch.float64).contiguous()
        idx, _, _ = batch.device_state(cams.device, analysis.layout)
        obs = st["observations"].detach().to(torch.float64).contiguous()
        out = torch.empty((batch.count, 2), dtype=torch.float64, device=cams.device)
        loss = batch.loss
        delta = loss.delta if isinstance(loss, HuberLoss) else -1.0
        if loss is not None and not isinstance(loss, HuberLoss):
            raise NotImplementedError("implicit gradients support the Huber loss or no loss")
        wts = batch.device_state(cams.device, analysis.layout)[2]
        lib = _lib.load()
        _lib.check(lib.tlh_reprojection_obs_vjp(
            cams.data_ptr(), pts.data_ptr(), obs.data_ptr(), idx[0].contiguous().data_ptr(),
            idx[1].contiguous().data_ptr(), w.contiguous().data_ptr(),
            analysis.layout[batch.type_indices[0]].offset, analysis.layout[batch.type_indices[1]].offset,
            float(delta), batch.count, out.data_ptr(), _lib.stream_handle(cams.device)), "tlh_reprojection_obs_vjp")
        if wts is not None:
            out = out * wts[:, None]
        st["adjoint"] = w
        return out.to(st["observations"].dtype).reshape(st["observations"].shape), None


def solve_implicit(problem, observations: torch.Tensor, *, adjoint_damping: float = 0.0, refine: int = 0,
                   polish: int = 0, adjoint_config: ConjugateGradientConfig | None = None, **options) -> dict:
    """Solve ``problem`` and return {name: tensor} differentiable w.r.t. ``observations``.

    ``observations`` must be the tensor given to ``add_residual`` for a
    reprojection batch.  The returned tensors carry a grad_fn whose backward
    runs the adjoint solve described in the module docstring.  The problem
    should be gauge-fixed (e.g. priors on one camera and one point) so that
    the optimum depends smoothly on the observations.  ``polish`` Newton steps on the
    exact Hessian (``newton_polish``) follow the LM solve, so that the solution and
    the adjoint are evaluated at the optimum to float64 rounding.
    """
    k, pos, opts = _find_batch(problem, observations)
    names = list(problem.parameters)
    state = dict(problem=problem, options=options, batch=k, arg=pos, opts=opts, observations=observations,
                 adjoint_damping=float(adjoint_damping), refine=int(refine), polish=int(polish),
                 adjoint_config=adjoint_config or ConjugateGradientConfig(rtol=1e-12, atol=0.0, max_iterations=2000,
                                                                         residual_check_interval=25))
    outs = _ImplicitSolve.apply(observations, state)
    result = dict(zip(state["names"], outs))
    problem.parameters.update(result)
    return result
