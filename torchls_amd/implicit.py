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
   the exact Hessian (``torch.func`` Hessian-vector products of the float64
   objective), preconditioned by that Gauss-Newton solve, which removes the
   Gauss-Newton approximation error (large for Huber outliers, whose exact
   curvature along their residual vanishes);
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


def _objective(analysis, values, delta):
    """float64 0.5 * sum rho(||r||^2) at retract(values, delta) (differentiable)."""
    moved = analysis.retract(values, delta)
    first = next(iter(moved.values()))
    total = torch.zeros((), dtype=torch.float64, device=first.device)
    for b in analysis.batches:
        r = b.residuals(moved, analysis.layout, first.device, first.dtype).to(torch.float64)
        s = r.square().sum(-1)
        v = s if b.loss is None else b.loss.value(s)
        w = b.device_state(first.device, analysis.layout)[2]
        if w is not None:
            v = v * w
        total = total + 0.5 * v.sum()
    return total


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
        lin = analysis.linearize(values)
        J = lin.jacobian
        lam = st["adjoint_damping"]
        op = HipNormalPlan(J, dtype=torch.float64).bind(J)
        diag = HipDiagonalPlan(J).diagonal_blocks(J)
        pre = HipBlockJacobi(J.layout, tuple(d.double() for d in diag), lam)
        cfg = st["adjoint_config"]
        matvec = lambda v: op.normal_matvec(v, lam)  # noqa: E731
        gn_solve = lambda v: hip_pcg(matvec, v, preconditioner=pre, config=cfg).solution  # noqa: E731
        w = gn_solve(b)
        if st["refine"] > 0:
            zero = torch.zeros_like(w)

            def hvp(v):
                with torch.enable_grad():
                    hv = torch.func.jvp(torch.func.grad(lambda d: _objective(analysis, values, d)), (zero,), (v,))[1]
                return hv + lam * v

            w = _exact_adjoint(hvp, gn_solve, b, w, st["refine"])
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
                   adjoint_config: ConjugateGradientConfig | None = None, **options) -> dict:
    """Solve ``problem`` and return {name: tensor} differentiable w.r.t. ``observations``.

    ``observations`` must be the tensor given to ``add_residual`` for a
    reprojection batch.  The returned tensors carry a grad_fn whose backward
    runs the adjoint solve described in the module docstring.  The problem
    should be gauge-fixed (e.g. priors on one camera and one point) so that
    the optimum depends smoothly on the observations.
    """
    k, pos, opts = _find_batch(problem, observations)
    names = list(problem.parameters)
    state = dict(problem=problem, options=options, batch=k, arg=pos, opts=opts, observations=observations,
                 adjoint_damping=float(adjoint_damping), refine=int(refine),
                 adjoint_config=adjoint_config or ConjugateGradientConfig(rtol=1e-12, atol=0.0, max_iterations=2000,
                                                                         residual_check_interval=25))
    outs = _ImplicitSolve.apply(observations, state)
    result = dict(zip(state["names"], outs))
    problem.parameters.update(result)
    return result
