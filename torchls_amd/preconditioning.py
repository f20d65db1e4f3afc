"""Joint block-Jacobi groups on the HIP backend (drop-in for preconditioning.py:20-134).

``ConjugateGradientConfig.preconditioner_groups`` joins variable groups with
identical ids row-wise: variable v of every member forms one block of the
summed member dims, holding the members' diagonal blocks and the cross terms
J_a^T J_b of residuals whose slots of two member types refer to the same id
(preconditioning.py:47-121).  The blocks are inverted in float64
(tlh_block_inverse, up to 16 x 16) and applied by tlh_block_apply on the
gathered member columns.  The cross terms are summed per variable in a fixed
order (a stable sort by variable, tlh_segment_sum_f64), not with atomics.
"""

from __future__ import annotations

from dataclasses import dataclass

import torch

from . import _lib
from .hip_diagonal import LinAlgError
from .variables import Var


def normalize_groups(groups):
    """Validate joint groups (preconditioning.py:20-37: the same rules and messages)."""
    if not isinstance(groups, (tuple, list)):
        raise ValueError("preconditioner_groups must be a sequence of groups")
    normalized, seen = [], set()
    for group in groups:
        if not isinstance(group, (tuple, list)) or len(group) < 2:
            raise ValueError("Each joint preconditioner group needs at least two variable keys")
        for key in group:
            if not ((isinstance(key, str) and key) or (isinstance(key, type) and issubclass(key, Var))):
                raise ValueError("Preconditioner keys must be parameter names or Var types")
            if key in seen:
                raise ValueError("A variable may occur in only one preconditioner group")
            seen.add(key)
        normalized.append(tuple(group))
    return tuple(normalized)


@dataclass(frozen=True)
class JacobiBlock:
    members: tuple
    count: int
    dim: int
    columns: torch.Tensor  # int64 [count, dim]: tangent columns of each block row


class BlockJacobiPlan:
    """Member groups joined row-wise; other layout entries stay separate blocks."""

    MAX_DIM = 16

    def __init__(self, layout, groups, device):
        self.layout = layout
        lookup = {}
        for i, entry in enumerate(layout):
            lookup[entry.key] = i
            if entry.name is not None:
                lookup[entry.name] = i
        grouped, selected = [], set()
        for group in normalize_groups(groups):
            if any(key not in lookup for key in group):
                raise ValueError("Unknown preconditioner key (Schur groups must be retained variables)")
            members = tuple(lookup[key] for key in group)
            first = layout[members[0]]
            if any(not torch.equal(first.ids.cpu(), layout[i].ids.cpu()) for i in members[1:]):
                raise ValueError("Joint preconditioner groups require identical IDs in identical order")
            grouped.append(members)
            selected.update(members)
        grouped.extend((i,) for i in range(len(layout)) if i not in selected)
        self.blocks, self.locations = [], {}
        for b, members in enumerate(grouped):
            offset, columns = 0, []
            for i in members:
                e = layout[i]
                self.locations[i] = (b, offset)
                columns.append(e.offset + torch.arange(e.count, device=device)[:, None] * e.dim
                               + torch.arange(e.dim, device=device)[None, :])
                offset += e.dim
            if offset > self.MAX_DIM:
                raise ValueError(f"joint preconditioner blocks are limited to {self.MAX_DIM} columns")
            self.blocks.append(JacobiBlock(members, layout[members[0]].count, offset,
                                           torch.cat(columns, dim=1).contiguous()))
        self.device = device
        self._cross = None

    def normal_blocks(self, jacobian, diagonal, *, col_scale: torch.Tensor | None = None):
        """float64 joint blocks: member diagonal blocks plus same-id cross terms.  With
        ``col_scale`` the cross terms are those of J diag(col_scale) (``diagonal`` is then
        expected scaled already, as the solver passes it)."""
        blocks = [torch.zeros((b.count, b.dim, b.dim), dtype=torch.float64, device=self.device) for b in self.blocks]
        for i, e in enumerate(self.layout):
            b, o = self.locations[i]
            blocks[b][:, o: o + e.dim, o: o + e.dim] = diagonal[i].to(torch.float64)
        if self._cross is None:
            self._cross = []
            for g, group in enumerate(jacobian.groups):
                for p, i in enumerate(group.type_indices):
                    bi, a = self.locations[i]
                    for q in range(p):
                        j = group.type_indices[q]
                        bj, c = self.locations[j]
                        # same-type aliases are already in the diagonal blocks
                        if bi == bj and i != j:
                            rows = group.variable_indices[p]
                            keep = torch.nonzero(rows == group.variable_indices[q]).reshape(-1)
                            if keep.numel():
                                var = rows[keep].to(torch.int64)
                                order = torch.sort(var, stable=True).indices
                                ptr = torch.zeros(self.blocks[bi].count + 1, dtype=torch.int64, device=self.device)
                                ptr[1:] = torch.cumsum(torch.bincount(var, minlength=self.blocks[bi].count), 0)
                                self._cross.append((g, p, q, bi, a, c, keep, order.to(torch.int32),
                                                    ptr.to(torch.int32)))
        lib = _lib.load()
        stream = _lib.stream_handle(self.device)
        for g, p, q, bi, a, c, keep, order, ptr in self._cross:
            group = jacobian.groups[g]
            left = group.blocks[p][keep].to(torch.float64)
            right = group.blocks[q][keep].to(torch.float64)
            if col_scale is not None:
                left = left * col_scale[group.columns[p][keep]][:, None, :]
                right = right * col_scale[group.columns[q][keep]][:, None, :]
            cross = torch.bmm(left.transpose(1, 2).contiguous(), right).reshape(keep.numel(), -1).contiguous()
            da, dc = left.shape[2], right.shape[2]
            count = self.blocks[bi].count
            total = torch.empty((count, da * dc), dtype=torch.float64, device=self.device)
            _lib.check(lib.tlh_segment_sum_f64(cross.data_ptr(), da * dc, 0, da * dc, order.data_ptr(), ptr.data_ptr(),
                                               count, total.data_ptr(), stream), "tlh_segment_sum_f64")
            total = total.reshape(count, da, dc)
            blocks[bi][:, a: a + da, c: c + dc] += total
            blocks[bi][:, c: c + dc, a: a + da] += total.transpose(1, 2)
        return blocks

    def factor(self, blocks, damping, *, graph_safe=False, defer_check=False):
        return HipJointBlockJacobi([(b.count, b.dim, b.columns) for b in self.blocks], blocks, damping,
                                   defer_check=defer_check)


class HipJointBlockJacobi:
    """Callable z = (B + lambda I)^-1 r per joint block (float64 inverses).  ``meta`` lists
    each block's (count, dim, columns [count, dim]) in the coordinates of the vectors it
    is applied to."""

    def __init__(self, meta, blocks, damping: float, *, status: torch.Tensor | None = None,
                 defer_check: bool = False):
        self.meta = [(int(c), int(d), cols.reshape(-1).contiguous()) for c, d, cols in meta]
        self.lib = _lib.load()
        dev = blocks[0].device
        stream = _lib.stream_handle(dev)
        self.status = torch.zeros(1, dtype=torch.int32, device=dev) if status is None else status
        self.deferred = bool(defer_check)
        self.inverses = []
        for (count, dim, _), value in zip(self.meta, blocks):
            value = value.to(torch.float64).contiguous()
            inv = torch.empty(max(1, count * dim * (dim + 1) // 2), dtype=torch.float64, device=dev)
            _lib.check(self.lib.tlh_block_inverse(value.data_ptr(), count, dim, float(damping), inv.data_ptr(),
                                                  self.status.data_ptr(), stream), "tlh_block_inverse")
            self.inverses.append(inv)
        if not self.deferred:
            self.check()

    def check(self) -> None:
        if int(self.status.item()) != 0:
            raise LinAlgError("damped joint block is not positive definite")

    def __call__(self, vector: torch.Tensor) -> torch.Tensor:
        out = torch.empty_like(vector)
        stream = _lib.stream_handle(vector.device)
        for (count, dim, cols), inv in zip(self.meta, self.inverses):
            gathered = vector[cols].contiguous()
            solved = torch.empty_like(gathered)
            _lib.check(self.lib.tlh_block_apply(inv.data_ptr(), dim, None, count, gathered.data_ptr(),
                                                solved.data_ptr(), _lib.dtype_code(vector.dtype), stream),
                       "tlh_block_apply")
            out[cols] = solved
        return out
