"""CPU execution of a tile plan's data flow (test infrastructure).

``emulate_normal_matvec`` runs the reduction a StructurePlan describes --
stage by rank, segment sums, dense per-workgroup accumulators, positional
partials, CSR finish, direct writes -- in numpy, exactly as the HIP kernels
do, so that plans can be validated against a dense J^T J x on the CPU.
"""

from __future__ import annotations

import numpy as np

from torchls_amd.plan import REDUCE_DENSE, StructurePlan


def _np(a):
    return None if a is None else a.detach().cpu().numpy()


def emulate_normal_matvec(plan: StructurePlan, packed: list, x: np.ndarray, damping: float):
    """Execute the plan's data flow on the CPU (numpy), exactly as the kernels do.

    ``packed[g]`` is the group's component-major array [components, count] in
    sorted order.      """
    types = plan.types
    parts = [np.zeros((t.partial_count, t.dim)) for t in types]
    direct_out = []
    for g, gp in zip(plan.groups, packed):
        offs = np.cumsum([0] + [s.dim for s in g.slots])
        R = g.rows
        xs = []
        for s in g.slots:
            t = types[s.type]
            xs.append(x[t.offset : t.offset + t.count * t.dim].reshape(t.count, t.dim)[_np(s.index)])
        q = np.zeros((g.count, R))
        for si, s in enumerate(g.slots):
            J = gp[R * offs[si] : R * offs[si + 1]].T.reshape(g.count, R, s.dim)
            q += np.einsum("nrd,nd->nr", J, xs[si])
        dense = {}
        for si, s in enumerate(g.slots):
            J = gp[R * offs[si] : R * offs[si + 1]].T.reshape(g.count, R, s.dim)
            cont = np.einsum("nrd,nr->nd", J, q)
            t = types[s.type]
            if s.mode == REDUCE_DENSE:
                dense[si] = np.zeros((g.nwg, t.count, t.dim))
            tile_start, seg_ptr, seg_start = _np(g.tile_start), _np(s.seg_ptr), _np(s.seg_start)
            seg_target, seg_var = _np(s.seg_target), _np(s.seg_var)
            for tile in range(g.ntiles):
                a, b = int(tile_start[tile]), int(tile_start[tile + 1])
                stage = np.zeros((b - a, s.dim))
                rank = np.arange(b - a) if s.rank is None else _np(s.rank)[a:b].astype(np.int64)
                stage[rank] = cont[a:b]
                k0, k1 = int(seg_ptr[tile]), int(seg_ptr[tile + 1])
                for k in range(k0, k1):
                    lo = int(seg_start[k])
                    hi = int(seg_start[k + 1]) if k + 1 < k1 else b - a
                    total = stage[lo:hi].sum(axis=0)
                    if s.mode == REDUCE_DENSE:
                        dense[si][tile // g.tiles_per_wg, seg_target[k]] += total
                    elif seg_target[k] < 0:
                        v = int(seg_var[k])
                        direct_out.append((t, v, total))
                    else:
                        parts[s.type][seg_target[k]] = total
            if s.mode == REDUCE_DENSE:
                for w in range(g.nwg):
                    base = (s.dense_block + w) * t.count
                    parts[s.type][base : base + t.count] = dense[si][w]
    y = np.zeros_like(x)
    for ti, t in enumerate(types):
        csr_ptr, csr_idx = _np(t.csr_ptr), _np(t.csr_idx)
        acc = np.zeros((t.count, t.dim))
        for b in range(t.dense_blocks):
            acc += parts[ti][b * t.count : (b + 1) * t.count]
        for v in range(t.count):
            for k in range(csr_ptr[v], csr_ptr[v + 1]):
                acc[v] += parts[ti][csr_idx[k]]
        sl = slice(t.offset, t.offset + t.count * t.dim)
        y[sl] = acc.reshape(-1) + damping * x[sl]
    for t, v, total in direct_out:
        sl = slice(t.offset + v * t.dim, t.offset + (v + 1) * t.dim)
        y[sl] = total + damping * x[sl]
    return y


def pack_numpy(plan: StructurePlan, group_blocks: list) -> list:
    """CPU twin of tlh_pack_blocks: blocks [N, r, d] per slot -> [components, N] sorted."""
    out = []
    for g, blocks in zip(plan.groups, group_blocks):
        cols = [np.asarray(b)[_np(g.order)].reshape(g.count, -1).T for b in blocks]
        out.append(np.concatenate(cols, axis=0) if cols else np.zeros((0, g.count)))
    return out
