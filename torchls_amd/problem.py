"""Named-parameter modeling API and linearization (host side).

``Problem(parameters, manifolds=...)`` with ``add_residual(fn, ("name", idx),
data..., loss=, weights=, jacobian_fn=, jac_mode=, batch_size=, in_dims=)``
and ``solve(**options)`` keeps the interface of torchls.named.Problem
(named.py:24-154).  Layout offsets follow the parameter dict order
(named.py:108-114); global columns are ``index * dim + offset + arange(dim)``
(problem.py:138-148).

Linearization follows linearization.py:62-128: per residual batch,
``vmap(jacrev|jacfwd)`` of the residual at delta = 0 through the retraction's
JVP chart (or a user ``jacobian_fn``), then the loss rescaling and the optional
weights.  Residual functions may also provide batched evaluators as attributes
(``fn.batched_residual(*args)`` and ``fn.batched_linearize(*args)``, see
``torchls_amd.residuals``), which skip vmap for the benchmark residuals.
The float64 objectives run in HIP (tlh_trial_objective); the reprojection
residual with SE(3) poses is linearized and trial-evaluated entirely in HIP
(tlh_reprojection_linearize / tlh_reprojection_trial, see ``hip_operands``).
"""

from __future__ import annotations

from collections.abc import Callable, Mapping
from dataclasses import dataclass

import torch
from torch import Tensor

from .losses import HuberLoss, Loss
from .sparse import BlockJacobian, JacobianGroup, VariableLayout
from .variables import EuclideanParameterization, Var


def _ids(ids) -> Tensor:
    out = torch.as_tensor(ids)
    if out.dtype not in (torch.int32, torch.int64) or out.ndim > 1:
        raise ValueError("Variable indices must be scalar or 1D integer tensors")
    return out.to(dtype=torch.long)


@dataclass(frozen=True)
class _Ref:
    name: str
    index: Tensor


@dataclass(frozen=True)
class Linearization:
    """Model residual, Jacobian, and (with losses/weights) raw residual and objective."""

    residual: Tensor
    jacobian: BlockJacobian
    raw_residual: Tensor | None = None
    objective: Tensor | None = None
    # 0-d device count of observations with a nonfinite residual or block entry
    nonfinite: Tensor | None = None


class _Batch:
    """One add_residual call: slots, indices and data, validated at analysis."""

    def __init__(self, fn, args, *, jac_mode, batch_size, in_dims, jacobian_fn, loss, weights, layout):
        self.fn, self.jacobian_fn, self.loss = fn, jacobian_fn, loss
        self.jac_mode, self.batch_size, self.in_dims = jac_mode, batch_size, in_dims
        self.args = args
        self.slots = tuple(i for i, a in enumerate(args) if isinstance(a, _Ref))
        if not self.slots:
            raise ValueError("Every residual must reference at least one parameter group")
        lookup = {e.name: t for t, e in enumerate(layout)}
        sizes = [args[i].index.numel() for i in self.slots if args[i].index.ndim == 1]
        self.batched = bool(sizes)
        self.count = max(sizes, default=1)
        if self.count == 0 or any(s not in (1, self.count) for s in sizes):
            raise ValueError("Variable batch lengths must match (or be one)")
        self.type_indices, self.indices = [], []
        for i in self.slots:
            ref = args[i]
            if ref.name not in lookup:
                raise ValueError(f"Undeclared parameter group {ref.name!r}")
            t = lookup[ref.name]
            idx = ref.index.reshape(-1).expand(self.count).contiguous()
            if bool((idx < 0).any()) or bool((idx >= layout[t].count).any()):
                raise ValueError(f"Out-of-range indices for {ref.name!r}")
            self.type_indices.append(t)
            self.indices.append(idx)
        self.type_indices = tuple(self.type_indices)
        self.weights = None
        if weights is not None:
            w = torch.as_tensor(weights, dtype=torch.float64).detach()
            if w.ndim > 1 or w.numel() not in (1, self.count):
                raise ValueError("Weights must be a real scalar or one value per observation")
            if not bool((torch.isfinite(w) & (w >= 0)).all()):
                raise ValueError("Weights must be finite and nonnegative")
            self.weights = w.reshape(-1).clone()
        for i, a in enumerate(args):
            if i in self.slots:
                continue
            t = torch.as_tensor(a)
            if self._axis(i, t) == 0 and (t.ndim == 0 or t.shape[0] not in (1, self.count)):
                raise ValueError("Data batch length must match observations; use in_dims=None for shared data")
        self._device = {}

    def _axis(self, i, t):
        if self.in_dims is not None:
            return self.in_dims[i]
        return 0 if self.batched and t.ndim > 0 else None

    def device_state(self, device, layout):
        if device not in self._device:
            idx = tuple(i.to(device) for i in self.indices)
            cols = tuple(
                i[:, None] * layout[t].dim + layout[t].offset + torch.arange(layout[t].dim, device=device)
                for i, t in zip(idx, self.type_indices)
            )
            w = None if self.weights is None else self.weights.to(device).expand(self.count).contiguous()
            self._device[device] = (idx, cols, w)
        return self._device[device]

    def gather(self, values, layout, device, dtype, start=0, end=None):
        idx, _, _ = self.device_state(device, layout)
        out, s = [], 0
        for i, a in enumerate(self.args):
            if isinstance(a, _Ref):
                out.append(values[a.name][idx[s][start:end]])
                s += 1
            else:
                t = torch.as_tensor(a)
                if t.is_floating_point() and not isinstance(a, Tensor):
                    t = torch.as_tensor(a, dtype=dtype)
                if self._axis(i, t) is None:
                    t = t.unsqueeze(0)
                t = t.expand(self.count, *t.shape[1:])[start:end]
                out.append(t.to(device=device, dtype=dtype if t.is_floating_point() else t.dtype))
        return out

    def hip_operands(self, values, layout, device, dtype):
        """(cams, pts, obs, huber_delta) when the HIP reprojection kernels apply, else None.

        They apply to the reprojection residual (``residuals.reprojection``)
        called as fn(("poses", idx), ("points", idx), observations[N, 2]) with
        SE(3) poses ([n, 4, 4] or [n, 3, 4]), Euclidean 3-vector points, no
        custom Jacobian, and no loss or the Huber loss."""
        if getattr(self.fn, "hip_residual", None) != "reprojection" or self.jacobian_fn is not None:
            return None
        if self.slots != (0, 1) or len(self.args) != 3 or not self.batched:
            return None
        if self.in_dims is not None and tuple(self.in_dims) != (0, 0, 0):
            return None
        if self.loss is not None and type(self.loss) is not HuberLoss:
            return None
        from .geometry import SE3Var

        tc, tp = self.type_indices
        if layout[tc].var_type is not SE3Var or not layout[tp].var_type.is_euclidean or layout[tp].dim != 3:
            return None
        cams, pts = values[self.args[0].name], values[self.args[1].name]
        if cams.shape[1:] not in ((4, 4), (3, 4)) or pts.shape[1:] != (3,):
            return None
        obs = self.args[2]
        if not isinstance(obs, Tensor) or obs.shape != (self.count, 2):
            return None
        if obs.device != device or obs.dtype != dtype or not obs.is_contiguous():
            obs = obs.to(device=device, dtype=dtype).contiguous()
        delta = self.loss.delta if self.loss is not None else -1.0
        return cams.contiguous(), pts.contiguous(), obs, float(delta)

    def hip_between_operands(self, values, layout, device, dtype):
        """(poses, measurements) when the HIP between-factor linearization applies, else None.

        It applies to ``residuals.between`` called as fn(("poses", i), ("poses", j),
        measurements[N, 4, 4]) on one SE(3) group, with no custom Jacobian, loss or weights."""
        if getattr(self.fn, "hip_residual", None) != "between" or self.jacobian_fn is not None:
            return None
        if self.slots != (0, 1) or len(self.args) != 3 or not self.batched:
            return None
        if self.in_dims is not None and tuple(self.in_dims) != (0, 0, 0):
            return None
        if self.loss is not None or self.weights is not None:
            return None
        from .geometry import SE3Var

        ti, tj = self.type_indices
        if layout[ti].var_type is not SE3Var or layout[tj].var_type is not SE3Var:
            return None
        if self.args[0].name != self.args[1].name:
            return None
        poses = values[self.args[0].name]
        if poses.shape[1:] not in ((4, 4), (3, 4)):
            return None
        meas = self.args[2]
        if not isinstance(meas, Tensor) or meas.shape[0] != self.count or meas.shape[1:] not in ((4, 4), (3, 4)):
            return None
        if meas.device != device or meas.dtype != dtype or not meas.is_contiguous():
            meas = meas.to(device=device, dtype=dtype).contiguous()
        return poses.contiguous(), meas

    def residuals(self, values, layout, device, dtype) -> Tensor:
        batched = getattr(self.fn, "batched_residual", None)
        chunk = self.batch_size if isinstance(self.batch_size, int) else None
        if chunk is None or self.count <= chunk:
            args = self.gather(values, layout, device, dtype)
            return (batched(*args) if batched else torch.func.vmap(self.fn)(*args)).reshape(self.count, -1)
        parts = []
        for s in range(0, self.count, chunk):
            args = self.gather(values, layout, device, dtype, s, min(s + chunk, self.count))
            parts.append((batched(*args) if batched else torch.func.vmap(self.fn)(*args)).reshape(len(args[0]), -1))
        return torch.cat(parts)

    def _linearizer(self, layout, dtype):
        specs = tuple(layout[t].var_type for t in self.type_indices)
        jvps = tuple(getattr(s, "retract_jvp_fn", None) for s in specs)
        slots = self.slots

        def local(deltas, *args):
            upd = list(args)
            for pos, spec, jvp, d in zip(slots, specs, jvps, deltas):
                upd[pos] = args[pos] + jvp(args[pos], d) if jvp is not None else spec.retract_fn(args[pos], d)
            r = self.fn(*upd).reshape(-1)
            return r, r

        mode = self.jac_mode
        if mode == "auto":
            mode = None  # decided on the first sample
        return specs, local, mode

    def linearize(self, values, layout, device, dtype):
        """Raw residual [N, r] and raw blocks (tuple of [N, r, d]) for this batch."""
        batched = getattr(self.fn, "batched_linearize", None)
        chunk = self.batch_size if isinstance(self.batch_size, int) else None
        ranges = [(0, self.count)] if chunk is None else [(s, min(s + chunk, self.count))
                                                           for s in range(0, self.count, chunk)]
        res_parts, block_parts = [], []
        for s, e in ranges:
            args = self.gather(values, layout, device, dtype, s, e)
            if batched is not None:
                r, blocks = batched(*args)
            elif self.jacobian_fn is not None:
                blocks = torch.func.vmap(self.jacobian_fn)(*args)
                r = torch.func.vmap(self.fn)(*args).reshape(e - s, -1)
                if not isinstance(blocks, tuple) or len(blocks) != len(self.slots):
                    raise ValueError("Custom Jacobian must return one block per parameter argument")
            else:
                specs, local, mode = self._linearizer(layout, dtype)
                if mode is None:
                    sample = self.fn(*(a[0] for a in args))
                    mode = "reverse" if sample.numel() <= sum(sp.tangent_dim for sp in specs) else "forward"
                diff = torch.func.jacrev if mode == "reverse" else torch.func.jacfwd
                zeros = tuple(torch.zeros(sp.tangent_dim, dtype=dtype, device=device) for sp in specs)
                blocks, r = torch.func.vmap(diff(local, argnums=0, has_aux=True),
                                            in_dims=(None, *([0] * len(args))))(zeros, *args)
                r = r.reshape(e - s, -1)
            for b, t in zip(blocks, self.type_indices):
                if b.shape != (e - s, r.shape[1], layout[t].dim):
                    raise ValueError("Incorrect local Jacobian shape")
            res_parts.append(r)
            block_parts.append(tuple(blocks))
        if len(ranges) == 1:
            return res_parts[0], block_parts[0]
        return torch.cat(res_parts), tuple(torch.cat(p) for p in zip(*block_parts))


class Analysis:
    """Layout + residual batches of a Problem (AnalyzedLeastSquaresProblem's role)."""

    def __init__(self, layout, batches, tangent_dim):
        self.layout = layout
        self.batches = batches
        self.tangent_dim = tangent_dim
        self.has_loss = any(b.loss is not None or b.weights is not None for b in batches)
        self._red = {}
        self.plans = {}  # symbolic HIP plans, reused by every solve of this structure

    def _reduction(self, device):
        from .hip_ops import Reduction

        if device not in self._red:
            self._red[device] = Reduction(device, max(16, 2 * len(self.batches) + 2))
        return self._red[device]

    def residuals(self, values) -> Tensor:
        dev, dt = _values_device_dtype(values)
        return torch.cat([b.residuals(values, self.layout, dev, dt).reshape(-1) for b in self.batches])

    def _objective_terms(self, values, before_rows=None):
        """Per batch: HIP float64 (objective, reduction) sums; before_rows per batch or None."""
        from .hip_ops import trial_objective

        from .hip_ops import reprojection_trial

        dev, dt = _values_device_dtype(values)
        red = self._reduction(dev)
        for k, b in enumerate(self.batches):
            before = None if before_rows is None else before_rows[k]
            idx, _, w = b.device_state(dev, self.layout)
            hip = b.hip_operands(values, self.layout, dev, dt)
            if hip is not None:
                reprojection_trial(red, k, hip, idx, w, before)
                continue
            after = b.residuals(values, self.layout, dev, dt)
            if b.loss is None or isinstance(b.loss, HuberLoss):
                trial_objective(red, k, before, after, b.count, after.shape[1],
                                huber_delta=None if b.loss is None else b.loss.delta, weights=w)
            else:  # a user-defined Loss: its own tensor code, float64
                a64 = after.to(torch.float64)
                sq_new = a64.square().sum(-1)
                val = b.loss.value(sq_new)
                if before is None:
                    chg = torch.zeros_like(val)
                else:
                    b64 = before.to(torch.float64)
                    chg = b.loss._reduction(b64.square().sum(-1), sq_new, ((b64 - a64) * (b64 + a64)).sum(-1))
                if w is not None:
                    val, chg = val * w, chg * w
                red.out[k, 0] = val.sum()
                red.out[k, 1] = chg.sum()
        if len(self.batches) == 1:
            return 0.5 * red.out[0]
        return 0.5 * red.out[: len(self.batches)].sum(0)

    def cost(self, values) -> Tensor:
        return self._objective_terms(values)[0]

    def evaluate_trial(self, values, lin: Linearization):
        """(objective, reduction) at ``values`` relative to ``lin`` (problem.py:315-348)."""
        before = lin.raw_residual if self.has_loss else lin.residual
        parts, off = [], 0
        for b, g in zip(self.batches, lin.jacobian.groups):
            n, rr = g.blocks[0].shape[:2]
            parts.append(before[off : off + n * rr].reshape(n, rr))
            off += n * rr
        terms = self._objective_terms(values, parts)
        return terms[0], terms[1]

    def linearize(self, values, *, block_dtype: torch.dtype | None = None) -> Linearization:
        """Linearize every residual batch.  ``block_dtype=torch.float32`` under float64
        parameters stores the Jacobian blocks of the HIP reprojection batches in float32
        (mixed precision: residuals, objective and the gradient sums stay float64)."""
        from .hip_ops import between_linearize, reprojection_linearize

        dev, dt = _values_device_dtype(values)
        groups, residuals, raws, objectives, bad = [], [], [], [], []
        for k, b in enumerate(self.batches):
            idx, cols, w = b.device_state(dev, self.layout)
            btw = b.hip_between_operands(values, self.layout, dev, dt)
            if btw is not None:
                red = self._reduction(dev)
                r, blocks, obj = between_linearize(red, k, btw, idx)
                if self.has_loss:  # another batch has a loss: this one contributes r^2
                    raws.append(r.reshape(-1))
                    objectives.append(obj)
                bad.append(red.out[k, 1])
                residuals.append(r.reshape(-1))
                if block_dtype is not None:  # every group of one Jacobian shares the block dtype
                    blocks = tuple(x.to(block_dtype) for x in blocks)
                groups.append(JacobianGroup(blocks, cols, b.type_indices, idx))
                continue
            hip = b.hip_operands(values, self.layout, dev, dt)
            if hip is not None:
                red = self._reduction(dev)
                raw, r, blocks, obj = reprojection_linearize(red, k, hip, idx, w, self.has_loss,
                                                             block_dtype=block_dtype)
                if self.has_loss:
                    raws.append(raw.reshape(-1))
                    objectives.append(obj)
                bad.append(red.out[k, 1])
                residuals.append(r.reshape(-1))
                groups.append(JacobianGroup(blocks, cols, b.type_indices, idx))
                continue
            r, blocks = b.linearize(values, self.layout, dev, dt)
            if self.has_loss:
                raws.append(r.reshape(-1))
                sq = r.to(torch.float64).square().sum(-1)
                if b.loss is None:
                    obj = sq
                else:
                    ev = b.loss.evaluate(sq)
                    obj = ev[0]
                    r, blocks = b.loss._linearize(r, blocks, sq, ev)
                if w is not None:
                    scale = w.sqrt().to(r.dtype)
                    r = r * scale[:, None]
                    blocks = tuple(x * scale[:, None, None] for x in blocks)
                    obj = obj * w
                objectives.append(obj.sum())
            residuals.append(r.reshape(-1))
            blocks = tuple(x.contiguous() for x in blocks)
            bad.append(sum((~torch.isfinite(x)).sum() for x in (r, *blocks)).to(torch.float64))
            if block_dtype is not None:  # every group of one Jacobian shares the block dtype
                blocks = tuple(x.to(block_dtype) for x in blocks)
            groups.append(JacobianGroup(blocks, cols, b.type_indices, idx))
        # one batch (the common case): its tensors are used as they are, no concatenation copies
        one = len(residuals) == 1
        residual = residuals[0] if one else torch.cat(residuals)
        J = BlockJacobian(tuple(groups), self.layout, (residual.numel(), self.tangent_dim))
        return Linearization(
            residual, J,
            (raws[0] if one else torch.cat(raws)) if self.has_loss else None,
            0.5 * (objectives[0] if one else torch.stack(objectives).sum()) if self.has_loss else None,
            bad[0].clone() if one else torch.stack(bad).sum(),  # bad[0] views a reused reduction slot
        )

    def retract(self, values, delta: Tensor, *, kernel: bool = False):
        """Retract every parameter group by its slice of ``delta``.  With
        ``kernel=True`` (the solver's non-differentiable update) SE(3) poses go
        through tlh_se3_retract and Euclidean groups through one add."""
        from .geometry import SE3Var

        out = {}
        for e in self.layout:
            x = values[e.name]
            flat = delta[e.offset : e.offset + e.count * e.dim]
            if kernel and e.var_type is SE3Var and x.shape[1:] in ((4, 4), (3, 4)) and delta.dtype == torch.float64:
                from . import _lib

                src = x.contiguous()
                moved = torch.empty_like(src)
                _lib.check(_lib.load().tlh_se3_retract(
                    _lib.dtype_code(src.dtype), src.data_ptr(), src.shape[1], e.count, flat.contiguous().data_ptr(),
                    e.dim, moved.data_ptr(), _lib.stream_handle(src.device)), "tlh_se3_retract")
                out[e.name] = moved
                continue
            d = flat.reshape(e.count, e.dim).to(x.dtype)
            if kernel and e.var_type.is_euclidean:
                out[e.name] = x + d.reshape(x.shape)
                continue
            out[e.name] = torch.func.vmap(e.var_type.retract_fn)(x, d)
        return out


def _values_device_dtype(values):
    first = next(iter(values.values()))
    return first.device, first.dtype


class Problem:
    """Own parameter tensors by name and bind residual inputs with (name, index)."""

    def __init__(self, parameters: Mapping[str, Tensor], *, manifolds: Mapping[str, type] | None = None):
        self.parameters = dict(parameters)
        self.manifolds = dict(manifolds or {})
        self._calls = []
        self._analysis = None
        self._signature = None
        self.last_summary = None

    def add_residual(self, residual_fn: Callable, *args, jac_mode="auto", batch_size=None, in_dims=None,
                     jacobian_fn=None, loss: Loss | None = None, weights=None) -> "Problem":
        if not callable(residual_fn):
            raise TypeError("The first argument must be a residual function")
        if loss is not None and not isinstance(loss, Loss):
            raise TypeError("loss must be a Loss instance such as HuberLoss, or None")
        if jac_mode not in ("auto", "reverse", "forward"):
            raise ValueError("jac_mode must be auto, reverse or forward")
        if in_dims is not None and (len(in_dims) != len(args) or any(x not in (0, None) for x in in_dims)):
            raise ValueError("in_dims must contain one 0 or None per argument")
        bound = tuple(
            _Ref(a[0], _ids(a[1]).detach().cpu().clone())
            if isinstance(a, tuple) and len(a) == 2 and isinstance(a[0], str) else a
            for a in args
        )
        self._calls.append((residual_fn, bound, dict(jac_mode=jac_mode, batch_size=batch_size, in_dims=in_dims,
                                                     jacobian_fn=jacobian_fn, loss=loss, weights=weights)))
        self._analysis = None
        return self

    def _analyze(self) -> Analysis:
        if not self.parameters or not self._calls:
            raise ValueError("A problem needs parameters and residuals")
        if set(self.manifolds) - set(self.parameters):
            raise ValueError("Manifold names must refer to declared parameter groups")
        signature, specs = [], []
        for name, value in self.parameters.items():
            if not isinstance(name, str) or not name:
                raise ValueError("Parameter group names must be nonempty strings")
            if not isinstance(value, Tensor) or value.ndim < 1 or value.numel() == 0:
                raise ValueError(f"{name}: expected a nonempty tensor with a leading parameter axis")
            spec = self.manifolds.get(name) or EuclideanParameterization(tuple(value.shape[1:]))
            if not isinstance(spec, EuclideanParameterization) and not (isinstance(spec, type) and issubclass(spec, Var)):
                raise TypeError(f"{name}: manifold must be a Var parameterization such as SE3Var")
            specs.append(spec)
            signature.append((name, tuple(value.shape), spec))
        signature = tuple(signature)
        if self._analysis is None or signature != self._signature:
            layout, off = [], 0
            for (name, value), spec in zip(self.parameters.items(), specs):
                e = VariableLayout(spec, torch.arange(value.shape[0]), off, name)
                layout.append(e)
                off += e.count * e.dim
            layout = tuple(layout)
            batches = [_Batch(fn, args, layout=layout, **opts) for fn, args, opts in self._calls]
            self._analysis = Analysis(layout, batches, off)
            self._signature = signature
        self._validate(self.parameters)
        return self._analysis

    def _validate(self, values):
        first = next(iter(values.values()))
        for name, value in values.items():
            if value.dtype not in (torch.float32, torch.float64):
                raise ValueError("Use float32 or float64 for least squares")
            if value.dtype != first.dtype or value.device != first.device:
                raise ValueError("All parameters must have the same dtype and device")

    def analyze(self) -> "Problem":
        self._analyze()
        return self

    @property
    def tangent_dim(self) -> int:
        return self._analyze().tangent_dim

    def residuals(self, parameters=None):
        return self._analyze().residuals(parameters or self.parameters)

    def cost(self, parameters=None):
        return self._analyze().cost(parameters or self.parameters)

    def linearize(self, parameters=None) -> Linearization:
        return self._analyze().linearize(parameters or self.parameters)

    def solve(self, *, return_summary: bool = False, **options):
        from .solvers import solve

        analysis = self._analyze()
        solution, summary = solve(analysis, dict(self.parameters), return_summary=True, **options)
        self.parameters.update(solution)
        self.last_summary = summary
        return (dict(solution), summary) if return_summary else dict(solution)
