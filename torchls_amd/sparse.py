"""Block-Jacobian containers (layout contract of torchls.sparse, sparse.py:11-29).

``JacobianGroup`` holds, per slot, observation-major blocks ``[N, r, d]`` and
the global tangent columns ``[N, d]`` (``index * dim + offset + arange(dim)``,
problem.py:138-148).  ``BlockJacobian`` is the whole linearization.  Its
operators (matvec, rmatvec, normal_matvec, squared_matvec_norm,
diagonal_blocks) run the HIP kernels of this package; there is no CPU path.
"""

from __future__ import annotations

from dataclasses import dataclass

import torch
from torch import Tensor


@dataclass(frozen=True)
class VariableLayout:
    """One variable type in the tangent vector (variables.py:100-117)."""

    var_type: object
    ids: Tensor
    offset: int
    name: str | None = None

    @property
    def key(self):
        return self.var_type if self.name is None else self.name

    @property
    def count(self) -> int:
        return int(self.ids.numel())

    @property
    def dim(self) -> int:
        return int(self.var_type.tangent_dim)


@dataclass(frozen=True)
class JacobianGroup:
    """Local Jacobian blocks and global tangent-column indices, per slot."""

    blocks: tuple  # each [observations, residual_dim, slot_dim]
    columns: tuple  # each [observations, slot_dim]
    type_indices: tuple
    variable_indices: tuple

    @property
    def num_rows(self) -> int:
        return self.blocks[0].shape[0] * self.blocks[0].shape[1]


@dataclass(frozen=True)
class BlockJacobian:
    groups: tuple
    layout: tuple
    shape: tuple

    def _operator(self):
        from .hip_ops import jacobian_operator

        return jacobian_operator(self)

    def matvec(self, vector: Tensor) -> Tensor:
        if vector.shape != (self.shape[1],):
            raise ValueError("matvec input has incorrect shape")
        return self._operator().matvec(vector)

    def rmatvec(self, vector: Tensor) -> Tensor:
        if vector.shape != (self.shape[0],):
            raise ValueError("rmatvec input has incorrect shape")
        return self._operator().rmatvec(vector)

    def normal_matvec(self, vector: Tensor, damping: float = 0.0) -> Tensor:
        return self._operator().normal_matvec(vector, damping)

    def squared_matvec_norm(self, vector: Tensor, *, batch_size: int = 262144) -> Tensor:
        if vector.shape != (self.shape[1],):
            raise ValueError("matvec input has incorrect shape")
        if batch_size < 1:
            raise ValueError("batch_size must be positive")
        return self._operator().squared_matvec_norm(vector)

    def diagonal_blocks(self) -> tuple:
        return self._operator().diagonal_blocks()

    def to_dense(self) -> Tensor:
        """Diagnostic dense copy (sparse.py:123-143); never used by the solver."""
        proto = self.groups[0].blocks[0]
        dense = proto.new_zeros(self.shape)
        offset = 0
        for group in self.groups:
            count, rows, _ = group.blocks[0].shape
            row_ids = torch.arange(count * rows, device=dense.device).reshape(count, rows, 1) + offset
            for block, columns in zip(group.blocks, group.columns, strict=True):
                dense.index_put_(
                    (
                        row_ids.expand_as(block).reshape(-1),
                        columns[:, None, :].expand_as(block).reshape(-1),
                    ),
                    block.reshape(-1),
                    accumulate=True,
                )
            offset += count * rows
        return dense
