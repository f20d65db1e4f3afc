"""float64 numpy restatement of torchls.sparse (sparse.py:11-143).

A ``NpJacobian`` is the numpy twin of ``BlockJacobian``: per group, one
``[N, r, d]`` block per slot, the slot's variable type and its variable
indices; the global column of (slot, observation, c) is
``idx * dim + offset + c`` (problem.py:138-148).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class NpGroup:
    blocks: list  # [N, r, d_s] float arrays
    types: list  # type index per slot
    indices: list  # [N] int64 per slot

    @property
    def count(self):
        return self.blocks[0].shape[0]

    @property
    def rows(self):
        return self.blocks[0].shape[1]


@dataclass
class NpJacobian:
    groups: list
    layout: list  # (count, dim, offset) per variable type

    @property
    def n(self):
        return sum(c * d for c, d, _ in self.layout)

    @property
    def m(self):
        return sum(g.count * g.rows for g in self.groups)

    def columns(self, g: NpGroup, s: int) -> np.ndarray:
        count, dim, offset = self.layout[g.types[s]]
        return g.indices[s][:, None] * dim + offset + np.arange(dim)[None, :]

    # sparse.py:31-41
    def matvec(self, x: np.ndarray) -> np.ndarray:
        assert x.shape == (self.n,)
        rows = []
        for g in self.groups:
            q = np.zeros((g.count, g.rows), dtype=np.result_type(x, g.blocks[0]))
            for s, b in enumerate(g.blocks):
                q += np.einsum("nrd,nd->nr", b, x[self.columns(g, s)])
            rows.append(q.reshape(-1))
        return np.concatenate(rows) if rows else np.zeros(0)

    # sparse.py:43-55
    def rmatvec(self, y: np.ndarray) -> np.ndarray:
        assert y.shape == (self.m,)
        out = np.zeros(self.n, dtype=np.result_type(y, self.groups[0].blocks[0]))
        off = 0
        for g in self.groups:
            local = y[off : off + g.count * g.rows].reshape(g.count, g.rows)
            for s, b in enumerate(g.blocks):
                c = np.einsum("nrd,nr->nd", b, local)
                out += np.bincount(self.columns(g, s).ravel(), weights=c.ravel(), minlength=self.n)
            off += g.count * g.rows
        return out

    # sparse.py:83-84
    def normal_matvec(self, x: np.ndarray, damping: float = 0.0) -> np.ndarray:
        return self.rmatvec(self.matvec(x)) + damping * x

    # sparse.py:57-81 (float64 accumulation of the squared norm)
    def squared_matvec_norm(self, x: np.ndarray) -> float:
        q = self.matvec(x).astype(np.float64)
        return float(np.dot(q, q))

    # sparse.py:86-102
    def scale_columns(self, scale: np.ndarray) -> "NpJacobian":
        groups = [
            NpGroup(
                [b * scale[self.columns(g, s)][:, None, :] for s, b in enumerate(g.blocks)],
                list(g.types),
                list(g.indices),
            )
            for g in self.groups
        ]
        return NpJacobian(groups, list(self.layout))

    # sparse.py:104-121 -- aliased slots add both cross terms where indices agree
    def diagonal_blocks(self) -> list:
        out = [np.zeros((c, d, d)) for c, d, _ in self.layout]
        for g in self.groups:
            for s, (b, t, idx) in enumerate(zip(g.blocks, g.types, g.indices)):
                gram = np.einsum("nri,nrj->nij", b, b)
                for o in range(s):
                    if g.types[o] == t:
                        cross = np.einsum("nri,nrj->nij", b, g.blocks[o])
                        mask = (idx == g.indices[o])[:, None, None]
                        gram = gram + (cross + np.swapaxes(cross, 1, 2)) * mask
                np.add.at(out[t], idx, gram)
        return out

    # sparse.py:123-143
    def to_dense(self) -> np.ndarray:
        dense = np.zeros((self.m, self.n))
        off = 0
        for g in self.groups:
            rows = off + np.arange(g.count * g.rows).reshape(g.count, g.rows)
            for s, b in enumerate(g.blocks):
                cols = self.columns(g, s)
                np.add.at(
                    dense,
                    (
                        np.broadcast_to(rows[:, :, None], b.shape).ravel(),
                        np.broadcast_to(cols[:, None, :], b.shape).ravel(),
                    ),
                    b.ravel(),
                )
            off += g.count * g.rows
        return dense


def random_jacobian(rng, layout_counts_dims, groups_spec, dtype=np.float64, scale=1.0):
    """Random block structure for tests.

    ``layout_counts_dims``: [(count, dim), ...];  ``groups_spec``: [(N, r, [type, ...]), ...].
    Indices are uniform random (aliases happen naturally for repeated types).
    """
    layout, off = [], 0
    for c, d in layout_counts_dims:
        layout.append((c, d, off))
        off += c * d
    groups = []
    for N, r, types in groups_spec:
        blocks = [(rng.standard_normal((N, r, layout[t][1])) * scale).astype(dtype) for t in types]
        idx = [rng.integers(0, layout[t][0], size=N).astype(np.int64) for t in types]
        groups.append(NpGroup(blocks, list(types), idx))
    return NpJacobian(groups, layout)
