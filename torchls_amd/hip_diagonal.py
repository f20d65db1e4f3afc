"""HIP diagonal blocks and block-Jacobi: drop-ins for warp_diagonal.py and block_jacobi.

``HipDiagonalPlan(jacobian).diagonal_blocks(jacobian)`` mirrors
``WarpDiagonalPlan`` (warp_diagonal.py:70-179): one full symmetric
``[count, d, d]`` block per layout entry, in the Jacobian's dtype, including
aliased-slot cross terms.  Internally the blocks are accumulated in float64
(``diagonal_blocks_f64``), which is what the solver's float64 inversions use.

``HipBlockJacobi`` holds float64 inverses of ``D_t + lambda I`` (closed-form
Cholesky per block, tlh_block_inverse) and applies them as a callable, like the
``apply`` closure returned by ``block_jacobi`` (solvers.py:259-285); the HIP PCG
fuses its application into the step kernel.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from .hip_normal import DeviceStructure, check_columns, packed_blocks, plan_for
from .plan import major_eligible, major_order


class LinAlgError(RuntimeError):
    """A damped diagonal block is not positive definite (torch.linalg.LinAlgError analogue)."""


class HipDiagonalPlan:
    def __init__(self, jacobian, *, shards=128, hot_threshold=512, dtype: torch.dtype | None = None):
        proto = jacobian.groups[0].blocks[0]
        if proto.device.type != "cuda" or proto.dtype not in (torch.float32, torch.float64):
            raise ValueError("HIP diagonal blocks require GPU float32 or float64 blocks")
        if shards < 1 or hot_threshold < 1:
            raise ValueError("shards and hot_threshold must be positive")
        self.device, self.block_dtype = proto.device, proto.dtype
        self.dtype = proto.dtype if dtype is None else dtype
        self.lib = _lib.load()
        check_columns(jacobian)
        self.layout = tuple((e.count, e.dim) for e in jacobian.layout)
        self.structure = tuple(
            (tuple(g.type_indices), tuple(g.variable_indices), tuple(b.shape for b in g.blocks))
            for g in jacobian.groups
        )
        self.plan, self.dev = plan_for(jacobian, self.dtype, allow_dense=False)
        self.major = self._major_slots(jacobian)
        self.partials = [
            torch.empty(max(1, t.sparse_partials * t.dim * (t.dim + 1) // 2), dtype=torch.float64,
                        device=self.device)
            for t in self.plan.types
        ]

    def _major_slots(self, jacobian) -> dict:
        """Types whose diagonal blocks come from tlh_major_gram: a type with few
        variables and many observations (cameras) seen by exactly one residual
        group, a two-slot 2x6 / 2x3 pair of different types.  type -> (group,
        slot, order, ptr)."""
        groups_of = {}
        for gi, g in enumerate(jacobian.groups):
            for t in set(g.type_indices):
                groups_of.setdefault(t, []).append(gi)
        out = {}
        for gi, g in enumerate(jacobian.groups):
            shape = [tuple(int(v) for v in b.shape[1:]) for b in g.blocks]
            if len(g.blocks) != 2 or g.type_indices[0] == g.type_indices[1] or \
                    sorted(shape) != [(2, 3), (2, 6)]:
                continue
            n = int(g.blocks[0].shape[0])
            for s, t in enumerate(g.type_indices):
                count = jacobian.layout[t].count
                if groups_of[t] != [gi] or not major_eligible(count, n):
                    continue
                order, ptr = major_order(g.variable_indices[s].detach().reshape(-1).to(self.device), count)
                out[t] = (gi, s, order, ptr)
        return out

    def compatible_with(self, normal_plan) -> bool:
        """True when a HipNormalPlan packs observations in this plan's order and tiles."""
        a, b = self.plan.groups, normal_plan.plan.groups
        return len(a) == len(b) and all(
            x.count == y.count and torch.equal(x.order, y.order) and torch.equal(x.tile_start, y.tile_start)
            for x, y in zip(a, b)
        )

    def _check(self, jacobian):
        if tuple((e.count, e.dim) for e in jacobian.layout) != self.layout or len(jacobian.groups) != len(
            self.structure
        ):
            raise ValueError("HIP diagonal plan requires unchanged incidence structure")
        for group, (types, indices, shapes) in zip(jacobian.groups, self.structure, strict=True):
            if (
                tuple(group.type_indices) != types
                or tuple(b.shape for b in group.blocks) != shapes
                or any(b.device != self.device or b.dtype != self.block_dtype for b in group.blocks)
                or len(group.variable_indices) != len(indices)
                or any(a is not b and not torch.equal(a, b) for a, b in zip(group.variable_indices, indices, strict=True))
            ):
                raise ValueError("HIP diagonal plan requires unchanged incidence structure, shapes and dtype")

    def _pack(self, jacobian) -> list:
        stream = _lib.stream_handle(self.device)
        packed = []
        for gi, (g, group) in enumerate(zip(self.plan.groups, jacobian.groups)):
            def pack(out, gi=gi, group=group):
                bufs = [None] * len(self.plan.groups)
                bufs[gi] = out
                op = self.dev.operator(self.dtype, bufs, [None] * len(self.plan.types))
                blocks = [b.contiguous() for b in group.blocks]
                ptrs = (ctypes.c_void_p * 3)(*[b.data_ptr() for b in blocks] + [None] * (3 - len(blocks)))
                _lib.check(self.lib.tlh_pack_blocks(ctypes.byref(op), gi, ptrs,
                                                    _lib.dtype_code(self.block_dtype), stream), "tlh_pack_blocks")

            packed.append(packed_blocks(jacobian, gi, self.dev.order_keys[gi], g.components, g.stride, self.dtype,
                                        pack))
        return packed

    def diagonal_blocks_f64(self, packed: list, jacobian=None) -> tuple:
        """float64 diagonal blocks from packed blocks laid out in this plan's order.

        With ``jacobian`` (the one ``packed`` was made from), the types in
        ``self.major`` are summed in variable order from its raw blocks instead."""
        outs = tuple(torch.empty((c, d, d), dtype=torch.float64, device=self.device) for c, d in self.layout)
        major = self.major if jacobian is not None else {}
        op = self.dev.operator(self.dtype, packed, self.partials)
        ptrs = (ctypes.c_void_p * len(outs))(*[None if t in major else o.data_ptr() for t, o in enumerate(outs)])
        lds = _lib.device_info(self.device).lds_per_block
        stream = _lib.stream_handle(self.device)
        _lib.check(self.lib.tlh_diagonal_blocks(ctypes.byref(op), ptrs, lds, stream), "tlh_diagonal_blocks")
        for t, (gi, s, order, ptr) in major.items():
            blocks = jacobian.groups[gi].blocks[s].contiguous()
            _, rows, dim = blocks.shape
            _lib.check(self.lib.tlh_major_gram(blocks.data_ptr(), _lib.dtype_code(blocks.dtype), rows, dim,
                                               order.data_ptr(), ptr.data_ptr(), outs[t].shape[0],
                                               outs[t].data_ptr(), stream), "tlh_major_gram")
        return outs

    def diagonal_blocks(self, jacobian) -> tuple:
        self._check(jacobian)
        outs = self.diagonal_blocks_f64(self._pack(jacobian), jacobian)
        return tuple(o.to(self.block_dtype) for o in outs)


class HipBlockJacobi:
    """Callable z = (D + lambda I)^-1 r per variable block, float64 inverses."""

    def __init__(self, layout, diagonal_f64, damping: float, *, status: torch.Tensor | None = None,
                 defer_check: bool = False):
        device = diagonal_f64[0].device
        self.lib = _lib.load()
        self.layout = [(e.count, e.dim, e.offset) for e in layout]
        self.size = sum(c * d for c, d, _ in self.layout)
        if status is None:
            status = torch.zeros(1, dtype=torch.int32, device=device)
        # defer_check: the positive-definiteness flag stays on the device and the
        # HIP PCG reads it together with its first scalar read (one sync less)
        self.status = status
        self.deferred = bool(defer_check)
        stream = _lib.stream_handle(device)
        self.inverses = []
        desc = _lib.BlockDiag()
        desc.ntypes = len(self.layout)
        for t, ((c, d, off), block) in enumerate(zip(self.layout, diagonal_f64)):
            block = block.to(torch.float64).contiguous()
            inv = torch.empty(max(1, c * d * (d + 1) // 2), dtype=torch.float64, device=device)
            _lib.check(self.lib.tlh_block_inverse(block.data_ptr(), c, d, float(damping), inv.data_ptr(),
                                                  status.data_ptr(), stream), "tlh_block_inverse")
            self.inverses.append(inv)
            desc.dim[t], desc.count[t], desc.offset[t], desc.inv[t] = d, c, off, inv.data_ptr()
        self.desc = desc
        self.device = device
        if not self.deferred:
            self.check()

    def check(self) -> None:
        """Raise LinAlgError if a damped block was not positive definite (one host sync)."""
        if int(self.status.item()) != 0:
            raise LinAlgError("damped diagonal block is not positive definite")

    def __call__(self, vector: torch.Tensor) -> torch.Tensor:
        from .hip_pcg import PcgState

        if vector.shape != (self.size,):
            raise ValueError("preconditioner input has incorrect shape")
        vector = vector.contiguous()
        r = torch.empty_like(vector)
        z = torch.empty_like(vector)
        p = torch.empty_like(vector)
        st = PcgState.scratch(vector.device)
        _lib.check(self.lib.tlh_pcg_reset(ctypes.byref(self.desc), vector.data_ptr(), None, r.data_ptr(),
                                          z.data_ptr(), p.data_ptr(), self.size, _lib.dtype_code(vector.dtype),
                                          ctypes.byref(st.desc), _lib.stream_handle(vector.device)),
                   "tlh_pcg_reset")
        return z
