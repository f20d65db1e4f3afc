"""HIP normal matvec: the drop-in for torchls.warp_normal (warp_normal.py:125-294).

``HipNormalPlan(jacobian)`` builds the symbolic tile plan once per solve;
``plan.bind(jacobian)`` packs one linearization's blocks into the plan's tiled
component-major layout; ``op.normal_matvec(v, damping)`` returns
``J^T (J v) + damping * v`` as a new tensor, like
``WarpNormalLinearization.normal_matvec``.  Reductions are deterministic.

``dtype`` selects the storage/arithmetic precision of the packed blocks and
vectors (default: the Jacobian's dtype).  ``dtype=torch.float32`` on a float64
Jacobian is the mixed-precision mode: float32 blocks and matvec, with the
caller keeping float64 outer quantities.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from .plan import REDUCE_DENSE, StructurePlan, build_structure


def _to_device(a: torch.Tensor, device) -> torch.Tensor:
    return a.to(device).contiguous()


def order_key(order: torch.Tensor) -> tuple:
    """Identity of a sorted order, computed on the device: its length and two wrapped
    int64 position-weighted sums (one host read)."""
    n = order.numel()
    if n == 0:
        return (0, 0, 0)
    o = order.to(torch.int64)
    w = torch.arange(n, dtype=torch.int64, device=o.device)
    s = torch.stack(((o * (2 * w + 1)).sum(), ((o + 1) * (o + 1) * (w + 7)).sum()))
    return (n, *s.tolist())


def _ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def check_columns(jacobian) -> None:
    """The tile plan addresses variables by index; columns must be index*dim+offset+arange."""
    for group in jacobian.groups:
        for cols, t, idx in zip(group.columns, group.type_indices, group.variable_indices, strict=True):
            entry = jacobian.layout[t]
            expect = idx.reshape(-1, 1).to(cols.dtype) * entry.dim + entry.offset
            expect = expect + torch.arange(entry.dim, device=cols.device, dtype=cols.dtype)
            if cols.shape != expect.shape or not torch.equal(cols, expect):
                raise ValueError("HIP operators require layout-ordered columns index*dim+offset+arange(dim)")


def packed_blocks(jacobian, group: int, order_key, components: int, stride: int, dtype, pack,
                  col_scale: torch.Tensor | None = None) -> torch.Tensor:
    """Packed [components, stride] blocks of one residual group, shared by every plan
    whose sorted order, stride, dtype and column scale match (``order_key``);
    ``pack(out)`` fills a new buffer.  Cached on the BlockJacobian, so each
    linearization packs once per scale."""
    cache = jacobian.__dict__.setdefault("_hip_packed", {})
    # in-place edits of the blocks (or of the scale) bump their version counters and force a re-pack
    blocks = jacobian.groups[group].blocks
    skey = None if col_scale is None else (col_scale.data_ptr(), col_scale._version)
    key = (group, order_key, components, stride, dtype, tuple((b.data_ptr(), b._version) for b in blocks), skey)
    out = cache.get(key)
    if out is None:
        out = torch.empty((components, max(stride, 1)), dtype=dtype, device=jacobian.groups[group].blocks[0].device)
        pack(out)
        cache[key] = out
    return out


class DeviceStructure:
    """Device copies of a StructurePlan plus the ctypes descriptors of the C ABI."""

    def __init__(self, plan: StructurePlan, device: torch.device):
        self.plan = plan
        self.device = device
        self.keep = []
        self._operators = {}  # (dtype, packed pointers, partial pointers) -> Operator
        # identity of each group's sorted order: plans with equal keys share packed blocks
        self.order_keys = [order_key(g.order) for g in plan.groups]
        put = lambda a: self._keep(_to_device(a, device))  # noqa: E731
        self.groups = (_lib.Group * max(1, len(plan.groups)))()
        for gi, g in enumerate(plan.groups):
            d = self.groups[gi]
            d.count = g.count
            d.rows = g.rows
            d.nslots = len(g.slots)
            d.ntiles = g.ntiles
            d.tile_cap = g.tile_cap
            d.nwg = g.nwg
            d.tiles_per_wg = g.tiles_per_wg
            d.stride = g.stride
            d.tile_start = put(g.tile_start).data_ptr()
            d.order = put(g.order).data_ptr()
            d.packed = None
            d.kernel = g.kernel
            d.primary = g.primary
            d.out_entries = g.out_entries
            d.stage_entries = g.stage_entries
            for si, s in enumerate(g.slots):
                ds = d.slot[si]
                ds.type = s.type
                ds.dim = s.dim
                ds.mode = s.mode
                ds.dense_block = s.dense_block
                # ranks are below the tile size (<= 2,048): stored as 16-bit values
                index, rank = s.index, None if s.rank is None else s.rank.to(torch.int16)
                if g.kernel == 1:  # wave groups read ids and ranks of whole tiles: pad to ntiles * tile_cap
                    pad = g.ntiles * g.tile_cap - g.count
                    index = torch.cat((index, index[-1:].expand(pad)))
                    rank = None if rank is None else torch.cat((rank, rank.new_zeros(pad)))
                ds.index = put(index).data_ptr()
                ds.rank = None if rank is None else put(rank).data_ptr()
                ds.seg_ptr = put(s.seg_ptr).data_ptr()
                ds.seg_start = put(s.seg_start).data_ptr()
                ds.seg_target = put(s.seg_target).data_ptr()
                ds.pos_base = s.pos_base
                ds.rounds = None if s.rounds is None else put(s.rounds.to(torch.int32)).data_ptr()
        self.types = (_lib.Type * len(plan.types))()
        for ti, t in enumerate(plan.types):
            d = self.types[ti]
            d.count = t.count
            d.dim = t.dim
            d.dense_blocks = t.dense_blocks
            d.offset = t.offset
            d.sparse_partials = t.sparse_partials
            d.csr_ptr = put(t.csr_ptr).data_ptr()
            d.csr_idx = put(t.csr_idx if t.csr_idx.numel() else torch.zeros(1, dtype=torch.int32)).data_ptr()
            d.partials = None
            d.direct = None if t.direct is None else put(t.direct).data_ptr()
            d.finish_list = None if t.finish_list is None else put(
                t.finish_list if t.finish_list.numel() else torch.zeros(1, dtype=torch.int32)).data_ptr()
            d.finish_count = t.count if t.finish_list is None else int(t.finish_list.numel())
            d.csr_entries = int(t.csr_idx.numel())
        self.sorted_index = None

    def _keep(self, t: torch.Tensor) -> torch.Tensor:
        self.keep.append(t)
        return t

    def operator(self, dtype: torch.dtype, packed: list, partials: list) -> "_lib.Operator":
        """Operator descriptor bound to packed blocks and partial workspaces.

        Cached per (dtype, block and partial addresses): the descriptor holds only
        those addresses and this structure's arrays, which are fixed after
        construction, and every library entry point reads it as const."""
        key = (dtype, tuple(_ptr(p) for p in packed), tuple(_ptr(p) for p in partials))
        op = self._operators.get(key)
        if op is not None:
            return op
        groups = (_lib.Group * max(1, len(self.plan.groups)))()
        ctypes.memmove(groups, self.groups, ctypes.sizeof(self.groups))
        for gi, p in enumerate(packed):
            groups[gi].packed = _ptr(p)
        types = (_lib.Type * len(self.plan.types))()
        ctypes.memmove(types, self.types, ctypes.sizeof(self.types))
        for ti, p in enumerate(partials):
            types[ti].partials = _ptr(p)
        op = _lib.Operator()
        op.dtype = _lib.dtype_code(dtype)
        op.ngroups = len(self.plan.groups)
        op.ntypes = len(self.plan.types)
        op.items = 0
        op.size = self.plan.size
        op.groups = ctypes.cast(groups, ctypes.POINTER(_lib.Group))
        op.types = ctypes.cast(types, ctypes.POINTER(_lib.Type))
        op._keep = (groups, types)  # descriptor arrays must outlive the Operator
        if len(self._operators) >= 64:
            self._operators.clear()
        self._operators[key] = op
        return op


# (rows, dim slot 0, dim slot 1) with an instantiated wave kernel (csrc/normal.hip)
# (6x6 pose-graph blocks run on the segment kernel: their wave kernel would
# hold 2 x 72 block values per item in registers and spill)
WAVE_SHAPES = frozenset({(2, 6, 3), (2, 3, 6)})


def plan_for(jacobian, dtype: torch.dtype, *, allow_dense: bool = True, tile_cap: int | None = None,
             overflow: bool = True):
    """Build the StructurePlan and its device copy for a BlockJacobian."""
    device = jacobian.groups[0].blocks[0].device
    info = _lib.device_info(device)
    layout = [(e.count, e.dim, e.offset) for e in jacobian.layout]
    groups = [
        (
            int(g.blocks[0].shape[1]),
            list(g.type_indices),
            [i.detach().reshape(-1).to(device=device, dtype=torch.int64) for i in g.variable_indices],
        )
        for g in jacobian.groups
    ]
    esize = 4 if dtype == torch.float32 else 8
    plan = build_structure(
        layout,
        groups,
        esize=esize,
        lds_limit=int(info.lds_per_block),
        cu_count=int(info.cu_count),
        allow_dense=allow_dense,
        overflow=overflow,
        tile_cap=tile_cap,
        wave_shapes=WAVE_SHAPES if allow_dense else frozenset(),
    )
    return plan, DeviceStructure(plan, device)


class HipNormalPlan:
    """Fixed incidence structure; arbitrary residual sizes, 1-3 slots and aliases.

    Mirrors ``WarpNormalPlan(jacobian, *, shards, hot_threshold, max_packed_bytes)``
    (warp_normal.py:132-185): the same ValueErrors for non-GPU / non-float
    input and 32-bit column overflow.  ``shards``/``hot_threshold``/
    ``max_packed_bytes`` are accepted for signature compatibility; the tile
    plan has no atomics to shard and always packs.
    """

    def __init__(self, jacobian, *, shards=128, hot_threshold=512, max_packed_bytes=256 * 1024**2,
                 dtype: torch.dtype | None = None, allow_dense: bool = True, tile_cap: int | None = None,
                 overflow: bool = True):
        proto = jacobian.groups[0].blocks[0]
        if proto.device.type != "cuda" or proto.dtype not in (torch.float32, torch.float64):
            raise ValueError("HIP normal matvec requires GPU float32 or float64 blocks")
        if shards < 1 or hot_threshold < 1:
            raise ValueError("shards and hot_threshold must be positive")
        if isinstance(max_packed_bytes, bool) or not isinstance(max_packed_bytes, int) or max_packed_bytes < 0:
            raise ValueError("max_packed_bytes must be a nonnegative integer")
        if jacobian.shape[1] > 2**31 - 1:
            raise ValueError("HIP normal matvec requires 32-bit column indices")
        self.dtype = proto.dtype if dtype is None else dtype
        if self.dtype not in (torch.float32, torch.float64):
            raise ValueError("HIP normal matvec computes in float32 or float64")
        self.block_dtype = proto.dtype
        self.device = proto.device
        self.size = jacobian.shape[1]
        self.lib = _lib.load()
        check_columns(jacobian)
        self.structure = tuple(
            (tuple(b.shape for b in g.blocks), tuple(g.type_indices), tuple(g.variable_indices))
            for g in jacobian.groups
        )
        self.plan, self.dev = plan_for(jacobian, self.dtype, allow_dense=allow_dense, tile_cap=tile_cap,
                                       overflow=overflow)
        for gi in range(len(self.plan.groups)):
            op = self.dev.operator(self.dtype, [None] * len(self.plan.groups), [None] * len(self.plan.types))
            need = self.lib.tlh_group_lds_bytes(ctypes.byref(op), gi)
            if need < 0 or need > _lib.device_info(self.device).lds_per_block:
                raise ValueError("tile plan exceeds the device LDS limit")

    def check_structure(self, jacobian) -> None:
        if jacobian.shape[1] != self.size or len(jacobian.groups) != len(self.structure):
            raise ValueError("HIP normal plan requires unchanged incidence structure")
        for group, (shapes, types, indices) in zip(jacobian.groups, self.structure, strict=True):
            if tuple(b.shape for b in group.blocks) != shapes or any(
                b.dtype != self.block_dtype or b.device != self.device for b in group.blocks
            ):
                raise ValueError("HIP normal plan requires unchanged block shapes, dtype and device")
            if tuple(group.type_indices) != types or len(group.variable_indices) != len(indices) or any(
                new is not old and not torch.equal(new, old)
                for new, old in zip(group.variable_indices, indices, strict=True)
            ):
                raise ValueError("HIP normal plan requires unchanged incidence structure")

    def bind(self, jacobian, *, col_scale: torch.Tensor | None = None) -> "HipNormalLinearization":
        """Pack one linearization.  ``col_scale`` (float64 [n], optional) binds the
        column-scaled Jacobian J diag(col_scale) instead (sparse.py:86-102), scaled
        while packing: no scaled copy of J is made."""
        self.check_structure(jacobian)
        if col_scale is not None:
            if col_scale.shape != (self.size,) or col_scale.device != self.device:
                raise ValueError("col_scale must be a vector of the tangent dimension on the plan's device")
            col_scale = col_scale.to(torch.float64).contiguous()
        return HipNormalLinearization(self, jacobian, col_scale)


class HipNormalLinearization:
    """One linearization packed into the plan layout (WarpNormalLinearization)."""

    hip_graph_safe = True  # normal_matvec launches only HIP kernels on the current stream
    pcg_method = "normal_matvec"  # described by pcg_operator for native PCG batches

    def __init__(self, plan: HipNormalPlan, jacobian, col_scale: torch.Tensor | None = None):
        self.plan = plan
        device, dtype = plan.device, plan.dtype
        self.partials = [
            torch.empty(max(1, t.partial_count * t.dim), dtype=dtype, device=device) for t in plan.plan.types
        ]
        stream = _lib.stream_handle(device)
        in_code = _lib.dtype_code(plan.block_dtype)
        self.packed = []
        for gi, (g, group) in enumerate(zip(plan.plan.groups, jacobian.groups)):
            def pack(out, gi=gi, group=group):
                bufs = [None] * len(plan.plan.groups)
                bufs[gi] = out
                op = plan.dev.operator(dtype, bufs, self.partials)
                blocks = [b.contiguous() for b in group.blocks]
                ptrs = (ctypes.c_void_p * 3)(*[b.data_ptr() for b in blocks] + [None] * (3 - len(blocks)))
                if col_scale is None:
                    _lib.check(plan.lib.tlh_pack_blocks(ctypes.byref(op), gi, ptrs, in_code, stream),
                               "tlh_pack_blocks")
                else:
                    _lib.check(plan.lib.tlh_pack_blocks_scaled(ctypes.byref(op), gi, ptrs, in_code,
                                                               col_scale.data_ptr(), stream), "tlh_pack_blocks_scaled")

            self.packed.append(packed_blocks(jacobian, gi, plan.dev.order_keys[gi], g.components, g.stride, dtype,
                                             pack, col_scale))
        self._col_scale = col_scale  # kept alive while the packed blocks may be re-read
        self.op = plan.dev.operator(dtype, self.packed, self.partials)

    def _check_vector(self, vector: torch.Tensor) -> None:
        p = self.plan
        if vector.shape != (p.size,) or vector.device != p.device or vector.dtype != p.dtype:
            raise ValueError("Vector must match HIP normal operator dimensions, device and dtype")

    def normal_matvec_into(self, vector: torch.Tensor, out: torch.Tensor, damping: float = 0.0) -> torch.Tensor:
        """Write J^T J v + damping v into ``out`` (no allocation; graph-capturable)."""
        _lib.check(
            self.plan.lib.tlh_normal_matvec(
                ctypes.byref(self.op), vector.data_ptr(), out.data_ptr(), float(damping),
                _lib.stream_handle(vector.device),
            ),
            "tlh_normal_matvec",
        )
        return out

    def pcg_operator(self, damping: float = 0.0) -> "_lib.PcgOperator":
        """Descriptor of ``normal_matvec(., damping)`` for the native PCG batch
        (tlh_pcg_iterations)."""
        A = _lib.PcgOperator()
        A.kind = _lib.PCG_OP_NORMAL
        A.op = ctypes.pointer(self.op)
        A.damping = float(damping)
        return A

    def normal_matvec(self, vector: torch.Tensor, damping: float = 0.0) -> torch.Tensor:
        self._check_vector(vector)
        vector = vector.contiguous()
        return self.normal_matvec_into(vector, torch.empty_like(vector), damping)
