"""Shared test helpers: numpy oracle structures -> torch BlockJacobian."""

from __future__ import annotations

import numpy as np
import torch

from torchls_amd.sparse import BlockJacobian, JacobianGroup, VariableLayout


class _Dim:
    def __init__(self, d):
        self.tangent_dim = d


def to_block_jacobian(J, device, dtype=torch.float64) -> BlockJacobian:
    layout = tuple(
        VariableLayout(_Dim(d), torch.arange(c), off) for c, d, off in J.layout
    )
    groups = []
    for g in J.groups:
        blocks = tuple(torch.tensor(b, dtype=dtype, device=device) for b in g.blocks)
        idx = tuple(torch.tensor(i, dtype=torch.int64, device=device) for i in g.indices)
        cols = tuple(
            i[:, None] * J.layout[t][1] + J.layout[t][2] + torch.arange(J.layout[t][1], device=device)
            for i, t in zip(idx, g.types)
        )
        groups.append(JacobianGroup(blocks, cols, tuple(g.types), idx))
    return BlockJacobian(tuple(groups), layout, (J.m, J.n))


def ba_jacobian(problem, delta=None):
    """Oracle NpJacobian of a generated bundle-adjustment problem at its initial guess."""
    from oracle.ba import linearize_ba
    from oracle.sparse import NpGroup, NpJacobian

    r, Jc, Jp, raw, cost = linearize_ba(
        problem.cameras_init, problem.points_init, problem.cam_idx, problem.pt_idx,
        problem.observations, delta,
    )
    layout = [(problem.n_cameras, 6, 0), (problem.n_points, 3, 6 * problem.n_cameras)]
    J = NpJacobian([NpGroup([Jc, Jp], [0, 1], [problem.cam_idx, problem.pt_idx])], layout)
    return J, r


def rel_err(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
