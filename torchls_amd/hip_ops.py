"""Host wrappers of the HIP float64 reductions and the BlockJacobian operators.

``LinearizationKernels`` binds one incidence structure (built once per solve)
and exposes, for any linearization with that structure:

* ``gradient(J, residual)`` -- float64 ``J^T r`` (tlh_rmatvec_f64);
* ``squared_matvec_norm(J, x)`` -- float64 ``||J x||^2``;
* ``matvec(J, x)`` -- ``J x`` in original residual order;
* ``trial(before, after, ...)`` -- float64 objective and trial reduction.

``jacobian_operator(J)`` gives the ``BlockJacobian`` methods a cached
``LinearizationKernels`` plus normal/diagonal plans for that Jacobian.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from .hip_diagonal import HipDiagonalPlan
from .hip_normal import HipNormalPlan


class Reduction:
    """Partials/counter scratch for the fixed-order float64 grid reductions."""

    def __init__(self, device, slots: int = 16):
        self.partials = torch.zeros(2 * _lib.TLH_PCG_MAX_BLOCKS, dtype=torch.float64, device=device)
        self.counter = torch.zeros(4, dtype=torch.int32, device=device)
        self.out = torch.zeros((slots, 2), dtype=torch.float64, device=device)


def lm_decide(trial, state: torch.Tensor, trust) -> torch.Tensor:
    """Gain ratio, acceptance and damping update of one LM trial on the device
    (tlh_lm_decide).  ``trial``: the float64 device tensors (trial objective, actual
    reduction, ||J delta||^2, step statistics of vector_stats, a value finite exactly
    when the step is), or None for a trial without a usable step; ``state``: float64
    device [damping, rejection factor], updated in place; ``trust``: ``_lib.LmTrust``.
    Returns the float64 [9] decision (include/torchls_hip.h)."""
    out = torch.empty(9, dtype=torch.float64, device=state.device)
    if trial is None:
        tr = None
    else:
        for t in trial:
            if t.dtype != torch.float64 or t.device != state.device:
                raise ValueError("lm_decide expects float64 tensors on the state's device")
        tr = ctypes.byref(_lib.LmTrial(*(t.data_ptr() for t in trial)))
    lib = _lib.load()
    _lib.check(lib.tlh_lm_decide(tr, state.data_ptr(), ctypes.byref(trust), out.data_ptr(),
                                 _lib.stream_handle(state.device)), "tlh_lm_decide")
    return out


class LinearizationKernels:
    def __init__(self, jacobian, diag_plan: HipDiagonalPlan | None = None):
        self.lib = _lib.load()
        self.device = jacobian.groups[0].blocks[0].device
        self.layout = jacobian.layout
        self.size = jacobian.shape[1]
        self.diag_plan = diag_plan if diag_plan is not None else HipDiagonalPlan(jacobian)
        self.red = Reduction(self.device, max(16, len(jacobian.groups) + 1))
        self.grad_partials = [
            torch.empty(max(1, t.sparse_partials * t.dim), dtype=torch.float64, device=self.device)
            for t in self.diag_plan.plan.types
        ]

    def _raw(self, group):
        dims = (ctypes.c_int32 * 3)(*[b.shape[2] for b in group.blocks] + [0] * (3 - len(group.blocks)))
        offsets = (ctypes.c_int64 * 3)(
            *[self.layout[t].offset for t in group.type_indices] + [0] * (3 - len(group.blocks))
        )
        index = [i.contiguous() for i in group.variable_indices]
        blocks = [b.contiguous() for b in group.blocks]
        iptr = (ctypes.c_void_p * 3)(*[i.data_ptr() for i in index] + [None] * (3 - len(index)))
        bptr = (ctypes.c_void_p * 3)(*[b.data_ptr() for b in blocks] + [None] * (3 - len(blocks)))
        return dims, offsets, iptr, bptr, (index, blocks)

    def gradient(self, jacobian, residual: torch.Tensor) -> torch.Tensor:
        """float64 J^T residual (residual in original row order, any float dtype)."""
        residual = residual.contiguous()
        op = self.diag_plan.dev.operator(torch.float64, [None] * len(jacobian.groups), self.grad_partials)
        stream = _lib.stream_handle(self.device)
        lds = _lib.device_info(self.device).lds_per_block
        # types with few, heavily observed variables (cameras) are summed in variable order when
        # their group takes the two-slot float64 kernel (16-byte aligned rows), which can skip a slot
        slot_of = {(gi, s): t for t, (gi, s, _, _) in self.diag_plan.major.items()}
        used = {}
        offset = 0
        keep = []
        for gi, group in enumerate(jacobian.groups):
            count, rows = group.blocks[0].shape[:2]
            blocks = [b.contiguous() for b in group.blocks]
            keep.append(blocks)
            v = residual[offset : offset + count * rows]
            offset += count * rows
            # the two-slot kernel reads float64 residuals and float64 or float32 blocks (vector loads)
            pair = (len(blocks) == 2 and rows == 2 and v.dtype == torch.float64 and v.data_ptr() % 16 == 0
                    and all(b.dtype in (torch.float64, torch.float32) and b.dtype == blocks[0].dtype
                            and b.data_ptr() % 16 == 0 for b in blocks))
            ptrs = []
            for s, blk in enumerate(blocks):
                if pair and (gi, s) in slot_of:
                    used[slot_of[(gi, s)]] = (blk, v)
                    ptrs.append(None)
                else:
                    ptrs.append(blk.data_ptr())
            bptr = (ctypes.c_void_p * 3)(*ptrs + [None] * (3 - len(blocks)))
            _lib.check(
                self.lib.tlh_rmatvec_f64(ctypes.byref(op), gi, bptr, _lib.dtype_code(blocks[0].dtype), v.data_ptr(),
                                         _lib.dtype_code(v.dtype), lds, stream),
                "tlh_rmatvec_f64",
            )
        out = torch.empty(self.size, dtype=torch.float64, device=self.device)
        start = 0
        for t in sorted(used) + [len(self.layout)]:
            if start < t:
                _lib.check(self.lib.tlh_rmatvec_finish_range_f64(ctypes.byref(op), start, t, out.data_ptr(), stream),
                           "tlh_rmatvec_finish_range_f64")
            start = t + 1
        for t, (blk, v) in used.items():
            _, _, order, ptr = self.diag_plan.major[t]
            ent = self.layout[t]
            _lib.check(self.lib.tlh_major_rmatvec(blk.data_ptr(), _lib.dtype_code(blk.dtype), v.data_ptr(),
                                                  _lib.dtype_code(v.dtype), blk.shape[1], ent.dim, order.data_ptr(),
                                                  ptr.data_ptr(), ent.count, out[ent.offset:].data_ptr(), stream),
                       "tlh_major_rmatvec")
        return out

    def squared_matvec_norm(self, jacobian, x: torch.Tensor) -> torch.Tensor:
        """float64 ||J x||^2 as a 0-d device tensor."""
        x = x.contiguous()
        stream = _lib.stream_handle(self.device)
        for gi, group in enumerate(jacobian.groups):
            count, rows = group.blocks[0].shape[:2]
            dims, offsets, iptr, bptr, keep = self._raw(group)
            _lib.check(
                self.lib.tlh_squared_matvec_norm(
                    count, rows, len(group.blocks), dims, offsets, iptr, bptr,
                    _lib.dtype_code(group.blocks[0].dtype), x.data_ptr(), _lib.dtype_code(x.dtype),
                    self.red.out[gi].data_ptr(), self.red.partials.data_ptr(), self.red.counter.data_ptr(), stream,
                ),
                "tlh_squared_matvec_norm",
            )
        if len(jacobian.groups) == 1:
            return self.red.out[0, 0]  # a view, valid until the next call
        return self.red.out[: len(jacobian.groups), 0].sum()

    def vector_stats(self, a: torch.Tensor | None, b: torch.Tensor | None) -> torch.Tensor:
        """float64 device tensor (a.b, b.b, max|b|, max|a|) of two equal-length vectors
        (either may be None), one launch (tlh_vector_stats)."""
        ref = a if a is not None else b
        out = torch.empty(4, dtype=torch.float64, device=ref.device)
        a = None if a is None else a.contiguous()
        b = None if b is None else b.contiguous()
        _lib.check(self.lib.tlh_vector_stats(
            None if a is None else a.data_ptr(), 0 if a is None else _lib.dtype_code(a.dtype),
            None if b is None else b.data_ptr(), 0 if b is None else _lib.dtype_code(b.dtype), ref.numel(),
            out.data_ptr(), self.red.partials.data_ptr(), self.red.counter.data_ptr(),
            _lib.stream_handle(ref.device)), "tlh_vector_stats")
        return out

    def matvec(self, jacobian, x: torch.Tensor) -> torch.Tensor:
        x = x.contiguous()
        stream = _lib.stream_handle(self.device)
        dtype = jacobian.groups[0].blocks[0].dtype
        out = torch.empty(jacobian.shape[0], dtype=dtype, device=self.device)
        offset = 0
        for group in jacobian.groups:
            count, rows = group.blocks[0].shape[:2]
            dims, offsets, iptr, bptr, keep = self._raw(group)
            _lib.check(
                self.lib.tlh_block_matvec(
                    count, rows, len(group.blocks), dims, offsets, iptr, bptr, _lib.dtype_code(dtype),
                    x.data_ptr(), _lib.dtype_code(x.dtype), out[offset:].data_ptr(), _lib.dtype_code(dtype), stream,
                ),
                "tlh_block_matvec",
            )
            offset += count * rows
        return out


def trial_objective(red: Reduction, slot: int, before, after, count: int, rows: int, *,
                    huber_delta: float | None = None, weights: torch.Tensor | None = None) -> None:
    """red.out[slot] = (sum w rho(||after||^2), sum w [rho(||before||^2) - rho(||after||^2)])."""
    lib = _lib.load()
    after = after.contiguous()
    before = None if before is None else before.contiguous()
    w = None if weights is None else weights.to(torch.float64).contiguous()
    stride = 0 if (w is None or w.numel() == 1) else 1
    _lib.check(
        lib.tlh_trial_objective(
            None if before is None else before.data_ptr(), _lib.dtype_code(before.dtype) if before is not None else 0,
            after.data_ptr(), _lib.dtype_code(after.dtype), count, rows, 0 if huber_delta is None else 1,
            float(huber_delta or 0.0), None if w is None else w.data_ptr(), stride, red.out[slot].data_ptr(),
            red.partials.data_ptr(), red.counter.data_ptr(), _lib.stream_handle(after.device),
        ),
        "tlh_trial_objective",
    )


class JacobianOperator:
    """Cached HIP plans behind the BlockJacobian methods."""

    def __init__(self, jacobian):
        self.kernels = LinearizationKernels(jacobian)
        self.jacobian = jacobian
        self._normal = None

    def matvec(self, x):
        return self.kernels.matvec(self.jacobian, x)

    def rmatvec(self, v):
        return self.kernels.gradient(self.jacobian, v).to(v.dtype)

    def normal_matvec(self, x, damping=0.0):
        if self._normal is None:
            self._normal = HipNormalPlan(self.jacobian).bind(self.jacobian)
        return self._normal.normal_matvec(x, damping)

    def squared_matvec_norm(self, x):
        return self.kernels.squared_matvec_norm(self.jacobian, x)

    def diagonal_blocks(self):
        return self.kernels.diag_plan.diagonal_blocks(self.jacobian)


def jacobian_operator(jacobian) -> JacobianOperator:
    proto = jacobian.groups[0].blocks[0]
    if proto.device.type != "cuda":
        raise ValueError("BlockJacobian operators run on the GPU (HIP); move the Jacobian to a cuda device")
    op = jacobian.__dict__.get("_hip_operator")
    if op is None:
        op = JacobianOperator(jacobian)
        object.__setattr__(jacobian, "_hip_operator", op)
    return op


def _reprojection_common(operands, idx, weights):
    cams, pts, obs, delta = operands
    n = idx[0].numel()
    wptr = None if weights is None else weights.data_ptr()
    return cams, pts, obs, delta, n, wptr


def reprojection_linearize(red: Reduction, slot: int, operands, idx, weights, want_raw: bool, block_dtype=None):
    """HIP linearization of one reprojection batch (tlh_reprojection_linearize).

    Returns (raw residual or None, model residual, (J_c, J_p), objective), the
    objective being the float64 sum w rho(||r||^2) as a 0-d view of
    ``red.out[slot, 0]``.  ``block_dtype`` (default: the parameters' dtype) may be
    float32 under float64 parameters: float32 blocks, float64 residuals."""
    cams, pts, obs, delta, n, wptr = _reprojection_common(operands, idx, weights)
    dt, dev = cams.dtype, cams.device
    bdt = dt if block_dtype is None else block_dtype
    raw = torch.empty((n, 2), dtype=dt, device=dev) if want_raw else None
    model = torch.empty((n, 2), dtype=dt, device=dev)
    jc = torch.empty((n, 2, 6), dtype=bdt, device=dev)
    jp = torch.empty((n, 2, 3), dtype=bdt, device=dev)
    lib = _lib.load()
    _lib.check(lib.tlh_reprojection_linearize(
        _lib.dtype_code(dt), _lib.dtype_code(bdt), cams.data_ptr(), cams[0].numel(), pts.data_ptr(), obs.data_ptr(),
        idx[0].data_ptr(), idx[1].data_ptr(), n, delta, wptr, None if raw is None else raw.data_ptr(),
        model.data_ptr(), jc.data_ptr(), jp.data_ptr(), red.out[slot].data_ptr(), red.partials.data_ptr(),
        red.counter.data_ptr(), _lib.stream_handle(dev)), "tlh_reprojection_linearize")
    return raw, model, (jc, jp), red.out[slot, 0]


def between_linearize(red: Reduction, slot: int, operands, idx):
    """HIP linearization of one SE(3) between-factor batch (tlh_between_linearize).

    Returns (residual [N, 6], (J_i, J_j) [N, 6, 6], objective) with the float64 sum of
    ||r||^2 as a 0-d view of ``red.out[slot, 0]`` (its nonfinite count in ``red.out[slot, 1]``)."""
    poses, meas = operands
    n = idx[0].numel()
    dt, dev = poses.dtype, poses.device
    res = torch.empty((n, 6), dtype=dt, device=dev)
    ji = torch.empty((n, 6, 6), dtype=dt, device=dev)
    jj = torch.empty((n, 6, 6), dtype=dt, device=dev)
    lib = _lib.load()
    _lib.check(lib.tlh_between_linearize(
        _lib.dtype_code(dt), poses.data_ptr(), poses[0].numel(), idx[0].data_ptr(), idx[1].data_ptr(),
        meas.data_ptr(), meas[0].numel(), n, res.data_ptr(), ji.data_ptr(), jj.data_ptr(), red.out[slot].data_ptr(),
        red.partials.data_ptr(), red.counter.data_ptr(), _lib.stream_handle(dev)), "tlh_between_linearize")
    return res, (ji, jj), red.out[slot, 0]


def reprojection_trial(red: Reduction, slot: int, operands, idx, weights, before) -> None:
    """red.out[slot] = (sum w rho(||r_new||^2), trial reduction) (tlh_reprojection_trial)."""
    cams, pts, obs, delta, n, wptr = _reprojection_common(operands, idx, weights)
    if before is not None:
        before = before.contiguous()
    lib = _lib.load()
    _lib.check(lib.tlh_reprojection_trial(
        _lib.dtype_code(cams.dtype), cams.data_ptr(), cams[0].numel(), pts.data_ptr(), obs.data_ptr(),
        idx[0].data_ptr(), idx[1].data_ptr(), n, delta, wptr,
        None if before is None else before.data_ptr(), _lib.dtype_code(before.dtype) if before is not None else 0,
        red.out[slot].data_ptr(), red.partials.data_ptr(), red.counter.data_ptr(), _lib.stream_handle(cams.device)),
        "tlh_reprojection_trial")
