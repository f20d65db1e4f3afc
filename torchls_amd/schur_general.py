"""Schur complement for any coupling structure (general path of the HIP backend).

The fused bundle-adjustment kernels (csrc/schur.hip) cover one residual group
that couples one retained slot with one eliminated slot.  The reference's
SchurPlan (schur.py:17-161) also allows a residual to couple several retained
slots with the eliminated variable (stereo pairs, camera plus intrinsics),
retained-retained coupling inside a residual, several coupling groups and
eliminated variables of more observations than a tile holds.  This module
serves those structures with the library's generic operators, without forming
W:

    S x  = J_R^T (J_R x - J_E C^-1 J_E^T J_R x) + lambda x      (schur.py:303-328)
    b_r' = b_r - J_R^T J_E C^-1 b_e                              (schur.py:285-290)
    x_e  = C^-1 (b_e - J_E^T J_R x)                              (schur.py:292-301)

with C = D_e + lambda I inverted per variable in float64 (tlh_block_inverse),
J x and J^T v by tlh_block_matvec / tlh_rmatvec_f64 (float64 accumulation),
C^-1 by tlh_block_apply.  The preconditioner blocks
D_r + lambda I - sum_e W_re C_e^-1 W_re^T (schur.py:330-378) sum each
(retained, eliminated) edge's W over its observations in a fixed order
(tlh_segment_sum_f64 over a stable sort by edge key), and the dense reduced
matrix (schur.py:380-397) is S applied to the identity, column by column.
"""

from __future__ import annotations

import torch

from . import _lib
from .hip_diagonal import HipBlockJacobi, LinAlgError


def _segment_sum(rows: torch.Tensor, order: torch.Tensor, ptr: torch.Tensor, count: int) -> torch.Tensor:
    """out[v] = sum of rows[order[k]] for k in [ptr[v], ptr[v+1]) (float64, fixed order)."""
    lib = _lib.load()
    rows = rows.reshape(rows.shape[0], -1).to(torch.float64).contiguous()
    dim = rows.shape[1]
    out = torch.empty((count, dim), dtype=torch.float64, device=rows.device)
    if count:
        _lib.check(lib.tlh_segment_sum_f64(rows.data_ptr(), dim, 0, dim, order.data_ptr(), ptr.data_ptr(), count,
                                           out.data_ptr(), _lib.stream_handle(rows.device)), "tlh_segment_sum_f64")
    return out


def _groups_by(key: torch.Tensor):
    """Stable sort by key: (order int32, unique keys, ptr int32 [U+1])."""
    order = torch.sort(key, stable=True).indices
    ukey, counts = torch.unique_consecutive(key[order], return_counts=True)
    ptr = torch.cat((counts.new_zeros(1), torch.cumsum(counts, 0)))
    return order.to(torch.int32).contiguous(), ukey, ptr.to(torch.int32).contiguous()


class GeneralSchurLinearization:
    """One linearization of a general Schur plan (``HipSchurPlan`` with ``general``)."""

    def __init__(self, plan, jacobian, diagonal, dense, col_scale=None):
        from .hip_ops import LinearizationKernels

        self.plan, self.jacobian, self.dense = plan, jacobian, dense
        self.diagonal = tuple(d.to(torch.float64).contiguous() for d in diagonal)
        if plan.kernels is None:
            plan.kernels = LinearizationKernels(jacobian)
        self.kernels = plan.kernels
        # col_scale arrives in the plan's (retained, eliminated) order; the operators
        # below work on full tangent vectors, so it is put back in layout order
        self.scale = None
        if col_scale is not None:
            full = torch.empty(plan.tangent_dim, dtype=torch.float64, device=plan.device)
            full[plan.global_indices] = col_scale[: plan.dim]
            full[plan.eoffset: plan.eoffset + plan.ecount * plan.edim] = col_scale[plan.dim:]
            self.scale = full

    def damped(self, damping: float, *, defer_check: bool = False) -> "GeneralSchurSystem":
        return GeneralSchurSystem(self, damping, defer_check=defer_check)


class GeneralSchurSystem:
    hip_graph_safe = False  # allocates its intermediate vectors through torch

    def __init__(self, lin: GeneralSchurLinearization, damping: float, *, defer_check: bool = False):
        self.lin, self.plan, self.damping = lin, lin.plan, float(damping)
        p = self.plan
        self.status = torch.zeros(1, dtype=torch.int32, device=p.device)
        self.deferred = bool(defer_check)
        d = p.edim
        self.cinv = torch.empty(max(1, p.ecount * d * (d + 1) // 2), dtype=torch.float64, device=p.device)
        _lib.check(p.lib.tlh_block_inverse(lin.diagonal[p.e].data_ptr(), p.ecount, d, self.damping,
                                           self.cinv.data_ptr(), self.status.data_ptr(),
                                           _lib.stream_handle(p.device)), "tlh_block_inverse")
        if not self.deferred:
            self.check()
        self._be = None

    def check(self) -> None:
        if int(self.status.item()) != 0:
            raise LinAlgError("damped eliminated block is not positive definite")

    # -- full-vector helpers (float64, layout order)
    def _esl(self):
        p = self.plan
        return slice(p.eoffset, p.eoffset + p.ecount * p.edim)

    def _J(self, v):
        """J (s * v) in residual order (float64)."""
        s = self.lin.scale
        return self.lin.kernels.matvec(self.lin.jacobian, v if s is None else v * s).to(torch.float64)

    def _Jt(self, u):
        """s * J^T u (float64 tangent vector)."""
        g = self.lin.kernels.gradient(self.lin.jacobian, u)
        return g if self.lin.scale is None else g * self.lin.scale

    def _cinv(self, z):
        p = self.plan
        z = z.to(torch.float64).contiguous()
        y = torch.empty_like(z)
        _lib.check(p.lib.tlh_block_apply(self.cinv.data_ptr(), p.edim, None, p.ecount, z.data_ptr(), y.data_ptr(),
                                         _lib.TLH_F64, _lib.stream_handle(p.device)), "tlh_block_apply")
        return y

    def _retained_full(self, x):
        p = self.plan
        full = torch.zeros(p.tangent_dim, dtype=torch.float64, device=p.device)
        full[p.global_indices] = x.to(torch.float64)
        return full

    def _eliminated_full(self, y):
        p = self.plan
        full = torch.zeros(p.tangent_dim, dtype=torch.float64, device=p.device)
        full[self._esl()] = y
        return full

    # -- DampedSchurSystem interface
    def reduced_rhs(self, rhs: torch.Tensor, *, scale: float = 1.0) -> torch.Tensor:
        p = self.plan
        b = scale * rhs.to(torch.float64)
        self._be = b[self._esl()].clone()
        t = self._Jt(self._J(self._eliminated_full(self._cinv(self._be))))
        return (b[p.global_indices] - t[p.global_indices]).to(p.dtype)

    def matvec(self, vector: torch.Tensor) -> torch.Tensor:
        p = self.plan
        if vector.shape != (p.dim,) or vector.dtype != p.dtype:
            raise ValueError("Schur matvec input must be the retained vector in the plan dtype")
        u = self._J(self._retained_full(vector))
        ye = self._cinv(self._Jt(u)[self._esl()])
        out = self._Jt(u - self._J(self._eliminated_full(ye)))
        return (out[p.global_indices] + self.damping * vector.to(torch.float64)).to(p.dtype)

    def backsubstitute(self, retained: torch.Tensor) -> torch.Tensor:
        p = self.plan
        if self._be is None:
            raise ValueError("reduced_rhs must be called before backsubstitute")
        full = self._retained_full(retained)
        t = self._Jt(self._J(full))[self._esl()]
        full[self._esl()] = self._cinv(self._be - t)
        return full

    def _cinv_full(self):
        """Full float64 C^-1 blocks [ecount, d, d] from the packed upper triangles."""
        p = self.plan
        d = p.edim
        tri = torch.triu_indices(d, d, device=p.device)
        cinv = torch.zeros((p.ecount, d, d), dtype=torch.float64, device=p.device)
        packed = self.cinv[: p.ecount * d * (d + 1) // 2].reshape(p.ecount, -1)
        cinv[:, tri[0], tri[1]] = packed
        cinv[:, tri[1], tri[0]] = packed
        return cinv

    def _edges(self, t):
        """(edge keys v * ecount + e sorted, W [E, dim_t, d]) of retained type t: each edge's
        W = sum over its observations of J_t^T J_e, in a fixed order."""
        p = self.plan
        J = self.lin.jacobian
        s = self.lin.scale
        rows, keys = [], []
        for g in J.groups:
            es = [q for q, ty in enumerate(g.type_indices) if ty == p.e]
            if not es:
                continue
            je = sum(g.blocks[q].to(torch.float64) for q in es)
            ie = g.variable_indices[es[0]]
            if s is not None:
                je = je * s[g.columns[es[0]]][:, None, :]
            for q, ty in enumerate(g.type_indices):
                if ty != t:
                    continue
                jr = g.blocks[q].to(torch.float64)
                if s is not None:
                    jr = jr * s[g.columns[q]][:, None, :]
                rows.append(torch.bmm(jr.transpose(1, 2).contiguous(), je))
                keys.append(g.variable_indices[q].to(torch.int64) * p.ecount + ie.to(torch.int64))
        if not rows:
            return None
        order, ukey, ptr = _groups_by(torch.cat(keys))
        W = _segment_sum(torch.cat(rows), order, ptr, ukey.numel()).reshape(-1, p.layout[t].dim, p.edim)
        return ukey, W

    def _per_variable(self, P, v, count):
        """Sum of P [E, a, b] over edges grouped by v (nondecreasing) -> [count, a, b]."""
        p = self.plan
        vptr = torch.zeros(count + 1, dtype=torch.int64, device=p.device)
        vptr[1:] = torch.cumsum(torch.bincount(v, minlength=count), 0)
        eorder = torch.arange(v.numel(), dtype=torch.int32, device=p.device)
        return _segment_sum(P, eorder, vptr.to(torch.int32), count).reshape(count, *P.shape[1:])

    def preconditioner_blocks(self):
        """float64 D_r + lambda I - sum over (retained, eliminated) edges W C^-1 W^T."""
        p = self.plan
        cinv = self._cinv_full()
        blocks = []
        for t in p.retained:
            ent = p.layout[t]
            M = self.lin.diagonal[t] + self.damping * torch.eye(ent.dim, dtype=torch.float64, device=p.device)
            edges = self._edges(t)
            if edges is not None:
                ukey, W = edges
                P = torch.bmm(torch.bmm(W, cinv[ukey % p.ecount]), W.transpose(1, 2).contiguous())
                M = M - self._per_variable(P, ukey // p.ecount, ent.count)
            blocks.append(M)
        return blocks

    def _joint_preconditioner(self):
        """Joint groups of retained variables (schur.py:331-352): each block holds the members'
        reduced diagonal blocks and, between two members a, b, the retained cross terms minus
        the sum over eliminated e linked to both of W_ae C_e^-1 W_be^T."""
        from .preconditioning import HipJointBlockJacobi

        p = self.plan
        jp = p.joint
        base = jp.normal_blocks(self.lin.jacobian, self.lin.diagonal, col_scale=self.lin.scale)
        cinv = self._cinv_full()
        edges = {t: self._edges(t) for t in p.retained}
        # retained coordinates of each full tangent column
        to_retained = torch.full((p.tangent_dim,), -1, dtype=torch.int64, device=p.device)
        to_retained[p.global_indices] = torch.arange(p.dim, device=p.device)
        meta, values = [], []
        for b, block in enumerate(jp.blocks):
            if p.e in block.members:
                continue
            M = base[b] + self.damping * torch.eye(block.dim, dtype=torch.float64, device=p.device)
            for ia, ta in enumerate(block.members):
                oa = jp.locations[ta][1]
                for tc in block.members[ia:]:
                    oc = jp.locations[tc][1]
                    ea, ec = edges[ta], edges[tc]
                    if ea is None or ec is None:
                        continue
                    ka, Wa = ea
                    kc, Wc = ec
                    pos = torch.searchsorted(ka, kc).clamp_max(ka.numel() - 1)
                    both = ka[pos] == kc
                    kk, Wa2, Wc2 = kc[both], Wa[pos[both]], Wc[both]
                    P = torch.bmm(torch.bmm(Wa2, cinv[kk % p.ecount]), Wc2.transpose(1, 2).contiguous())
                    S = self._per_variable(P, kk // p.ecount, block.count)
                    da, dc = Wa.shape[1], Wc.shape[1]
                    M[:, oa: oa + da, oc: oc + dc] -= S
                    if tc != ta:
                        M[:, oc: oc + dc, oa: oa + da] -= S.transpose(1, 2)
            meta.append((block.count, block.dim, to_retained[block.columns]))
            values.append(M)
        return HipJointBlockJacobi(meta, values, 0.0, status=self.status, defer_check=self.deferred)

    def preconditioner(self, *, graph_safe=False):
        from .sparse import VariableLayout

        p = self.plan
        if p.joint is not None:
            return self._joint_preconditioner()
        lay = tuple(VariableLayout(p.layout[t].var_type, p.layout[t].ids, off)
                    for t, (_, _, off) in zip(p.retained, p.rlayout[:-1]))
        return HipBlockJacobi(lay, self.preconditioner_blocks(), 0.0, status=self.status, defer_check=self.deferred)

    def dense_matrix(self) -> torch.Tensor:
        """The reduced matrix as S applied to each unit vector (symmetrised, float64)."""
        p = self.plan
        n = p.dim
        M = torch.empty((n, n), dtype=torch.float64, device=p.device)
        unit = torch.zeros(n, dtype=p.dtype, device=p.device)
        for j in range(n):
            unit.zero_()
            unit[j] = 1.0
            M[:, j] = self.matvec(unit).to(torch.float64)
        return 0.5 * (M + M.T)

    def dense_solve(self, rhs: torch.Tensor) -> torch.Tensor:
        from .cholesky import cholesky_solve_f64

        if self.deferred:
            self.check()
        return cholesky_solve_f64(self.dense_matrix(), rhs.to(torch.float64)).to(self.plan.dtype)
