"""Schur-complement linear solves on the HIP backend (drop-in for torchls.schur).

``HipSchurPlan(problem, eliminate, device, dtype)`` is the symbolic analysis
(SchurPlan, schur.py:17-161): the eliminated parameter group must be
independent per residual (schur.py:44-49); the retained groups keep their
layout order with re-based offsets (schur.py:27-37).  The fused HIP kernels
serve the bundle-adjustment structure: exactly one residual group ("W-group")
that couples one retained slot with one eliminated slot, any number of groups
on retained variables only (e.g. camera priors; one slot each for the dense
path), and any groups on eliminated variables only (e.g. point priors, which
only enter C = D_e + lambda I).  Every other structure the reference accepts
(several retained slots per residual, retained-retained coupling, several
coupling groups, eliminated variables with more observations than a tile) is
served by ``schur_general``.

``plan.linearize(J, diagonal)`` packs one linearization (SchurLinearization,
schur.py:192-257); ``.damped(lambda)`` gives the system with ``reduced_rhs``,
``matvec``, ``preconditioner``, ``backsubstitute`` and ``dense_solve``
(DampedSchurSystem, schur.py:263-397) -- all HIP kernels (csrc/schur.hip,
csrc/cholesky.hip); W is never stored.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from .hip_diagonal import LinAlgError
from .hip_normal import DeviceStructure, packed_blocks
from .plan import _stable_order, build_structure, major_edges, major_eligible


class HipSchurPlan:
    def __init__(self, problem, eliminate, device, *, dtype=torch.float64, preconditioner_groups=()):
        layout = problem.layout
        matches = [i for i, e in enumerate(layout) if e.key == eliminate or e.name == eliminate]
        if len(matches) != 1:
            raise ValueError("eliminate must identify one parameter group by name")
        self.e = matches[0]
        self.retained = tuple(i for i in range(len(layout)) if i != self.e)
        if not self.retained:
            raise ValueError("Schur elimination needs a retained parameter group")
        self.layout = layout
        self.device, self.dtype = device, dtype
        self.lib = _lib.load()
        ent = layout[self.e]
        self.ecount, self.edim, self.eoffset = ent.count, ent.dim, ent.offset
        # retained layout (re-based) then the eliminated type last
        self.local = {t: k for k, t in enumerate(self.retained)}
        self.local[self.e] = len(self.retained)
        rl, off = [], 0
        for t in self.retained:
            rl.append((layout[t].count, layout[t].dim, off))
            off += layout[t].count * layout[t].dim
        self.dim = off
        rl.append((ent.count, ent.dim, off))
        self.rlayout = rl
        gidx = [torch.arange(layout[t].offset, layout[t].offset + layout[t].count * layout[t].dim)
                for t in self.retained]
        self.global_indices = torch.cat(gidx).to(device)
        self.tangent_dim = sum(e.count * e.dim for e in layout)
        self.kernels = None  # LinearizationKernels of the general path
        self.wgroup = None
        self.groups = []  # (lin group index, [lin slot indices], [local types])
        # general: a structure outside the fused kernels' (one group coupling one retained
        # slot with one eliminated slot) -- served by schur_general
        self.general = False
        for gi, b in enumerate(problem.batches):
            es = [s for s, t in enumerate(b.type_indices) if t == self.e]
            rs = [s for s, t in enumerate(b.type_indices) if t != self.e]
            if len(es) > 1 and any(not torch.equal(b.indices[s], b.indices[es[0]]) for s in es[1:]):
                raise ValueError("Eliminated variables must be independent: a residual connects distinct eliminated indices")
            if es and rs:
                if self.wgroup is not None or len(es) != 1 or len(rs) != 1:
                    self.general = True
                    continue
                self.wgroup = len(self.groups)
                self.eslot = es[0]
                self.groups.append((gi, [0, 1], [self.local[b.type_indices[0]], self.local[b.type_indices[1]]]))
            elif rs:
                self.groups.append((gi, list(range(len(rs))), [self.local[b.type_indices[s]] for s in rs]))
        if self.wgroup is None:
            self.general = True
        # joint preconditioner groups of retained variables (schur.py:137-146, 331-352): served by
        # the general path's reduced preconditioner
        self.joint = None
        if preconditioner_groups:
            from .preconditioning import BlockJacobiPlan

            self.joint = BlockJacobiPlan(layout, preconditioner_groups, device)
            if any(self.e in b.members and len(b.members) > 1 for b in self.joint.blocks):
                raise ValueError("Unknown preconditioner key (Schur groups must be retained variables)")
            self.general = True
        self.retained_only_multi = any(len(sl) > 1 for k, (_, sl, _) in enumerate(self.groups) if k != self.wgroup)
        info = _lib.device_info(device)
        esize = 4 if dtype == torch.float32 else 8
        # the symbolic analysis runs on the device: index arrays are moved there once
        plan_groups = []
        for gi, slots, ltypes in self.groups:
            b = problem.batches[gi]
            plan_groups.append((None, ltypes, [b.indices[s].to(device=device, dtype=torch.int64) for s in slots]))
        self.rows = [None] * len(self.groups)
        self._plan_groups = plan_groups
        self._problem = problem
        self._info = info
        self._esize = esize
        self.plan = None
        self.sparse_plan = None
        self.pairs = None

    def _build(self, jacobian):
        """Plans need residual row counts, known from the first linearization."""
        if self.general:
            self.plan = True  # built: the general path has no tile plan
            return
        groups = []
        for k, (gi, slots, ltypes) in enumerate(self.groups):
            rows = int(jacobian.groups[gi].blocks[0].shape[1])
            self.rows[k] = rows
            groups.append((rows, ltypes, self._plan_groups[k][2]))
        eprim = {self.wgroup: self.eslot}
        # beyond the camera stage the plan reserves: the point stage and the wave tails
        threads = 1024 if self._esize == 4 else 512
        stage_extra = threads * self.edim * self._esize + threads // 64 * ((self.edim + 1) * 8 + 8)
        # the W-group kernel's workgroups (1,024 / 512 threads, csrc/schur.hip) fill a compute unit
        self.plan = build_structure(self.rlayout, groups, esize=self._esize,
                                    lds_limit=int(self._info.lds_per_block) - stage_extra,
                                    cu_count=int(self._info.cu_count), wgs_per_cu=1, force_primary=eprim,
                                    split_primary=False)
        # the preconditioner reduces W-group edge blocks only: its plan holds that group alone,
        # so every camera partial it lists is written by tlh_schur_precond
        self.sparse_plan = build_structure(self.rlayout, [groups[self.wgroup]], esize=self._esize,
                                           lds_limit=int(self._info.lds_per_block), allow_dense=False,
                                           force_primary={0: self.eslot}, split_primary=False)
        w = self.plan.groups[self.wgroup]
        prim = w.slots[self.eslot].index.to(torch.int64)
        cuts = w.tile_start[1:-1].to(torch.int64)
        if cuts.numel() and bool((prim[cuts] == prim[cuts - 1]).any()):
            # an eliminated variable has more observations than one tile holds
            self.general = True
            self.plan = True
            return
        # eliminated variables no coupling residual observes: back-substituted as C^-1 b_e alone
        seen = torch.zeros(self.ecount, dtype=torch.bool, device=self.device)
        seen[prim] = True
        lonely = torch.nonzero(~seen).reshape(-1)
        self.lonely = lonely.to(torch.int32) if lonely.numel() else None
        self.dev = DeviceStructure(self.plan, self.device)
        self.sdev = DeviceStructure(self.sparse_plan, self.device)
        # camera-order preconditioner (csrc/major.hip) for few, heavily observed retained variables
        self.major = None
        grp = jacobian.groups[self.groups[self.wgroup][0]]
        sr = 1 - self.eslot
        rcount = self.rlayout[self.groups[self.wgroup][2][sr]][0]
        if (tuple(grp.blocks[sr].shape[1:]) == (2, 6) and tuple(grp.blocks[self.eslot].shape[1:]) == (2, 3)
                and major_eligible(rcount, int(grp.blocks[0].shape[0]))):
            idx = self._plan_groups[self.wgroup][2]
            self.major = major_edges(idx[sr], idx[self.eslot], rcount)
        nt = len(self.retained)
        self.partials = [torch.empty(max(1, t.partial_count * t.dim), dtype=self.dtype, device=self.device)
                         for t in self.plan.types]
        self.pre_partials = [
            torch.empty(max(1, t.sparse_partials * t.dim * (t.dim + 1) // 2), dtype=torch.float64, device=self.device)
            for t in self.sparse_plan.types
        ]
        self.ntypes_ret = nt

    def _dense_pairs(self):
        """Observation pairs of every point run with cam(i) <= cam(j), grouped by camera pair
        (built on the device)."""
        if self.pairs is not None:
            return self.pairs
        w = self.plan.groups[self.wgroup]
        cs = 1 - self.eslot
        cam = w.slots[cs].index.to(torch.int64)
        prim = w.slots[self.eslot].index.to(torch.int64)
        n = prim.numel()
        head = torch.ones(n, dtype=torch.bool, device=self.device)
        head[1:] = prim[1:] != prim[:-1]
        starts = torch.nonzero(head).reshape(-1)
        lens = torch.diff(torch.cat((starts, starts.new_tensor([n]))))
        pi, pj = [], []
        for L in torch.unique(lens).tolist():
            s0 = starts[lens == L]
            a = torch.arange(L, dtype=torch.int64, device=self.device)
            ii = (s0[:, None, None] + a[None, :, None]).expand(-1, L, L).reshape(-1)
            jj = (s0[:, None, None] + a[None, None, :]).expand(-1, L, L).reshape(-1)
            keep = cam[ii] <= cam[jj]
            pi.append(ii[keep])
            pj.append(jj[keep])
        pi = torch.cat(pi)
        pj = torch.cat(pj)
        nc = int(self.rlayout[self.plan.groups[self.wgroup].slots[cs].type][0])
        key = cam[pi] * nc + cam[pj]
        o = _stable_order([key, pi, pj])
        pi, pj, key = pi[o], pj[o], key[o]
        ukey, counts = torch.unique_consecutive(key, return_counts=True)
        ptr = torch.cat((counts.new_zeros(1), torch.cumsum(counts, 0)))
        t = lambda a: a.to(torch.int32).contiguous()  # noqa: E731
        self.pairs = dict(ptr=t(ptr), i=t(pi), j=t(pj), ci=t(ukey // nc), cj=t(ukey % nc), n=int(ukey.numel()))
        return self.pairs

    def linearize(self, jacobian, diagonal, *, dense: bool = False, col_scale: torch.Tensor | None = None):
        """Pack one linearization; ``col_scale`` (float64, full tangent layout) makes
        the system that of J diag(col_scale), scaled while packing (``diagonal`` must
        then be the scaled blocks, as the reference passes them)."""
        if self.plan is None:
            self._build(jacobian)
        local_scale = None
        if col_scale is not None:
            total = sum(e.count * e.dim for e in self.layout)
            if col_scale.shape != (total,) or col_scale.device != self.device:
                raise ValueError("col_scale must be a vector of the tangent dimension on the plan's device")
            col_scale = col_scale.to(torch.float64)
            # the plan's layout: retained types (re-based), then the eliminated type
            local_scale = torch.cat((col_scale[self.global_indices],
                                     col_scale[self.eoffset: self.eoffset + self.ecount * self.edim])).contiguous()
        if self.general or (dense and self.retained_only_multi):
            from .schur_general import GeneralSchurLinearization

            return GeneralSchurLinearization(self, jacobian, diagonal, dense, local_scale)
        return HipSchurLinearization(self, jacobian, diagonal, dense, local_scale)


class HipSchurLinearization:
    def __init__(self, plan: HipSchurPlan, jacobian, diagonal, dense, col_scale=None):
        self.plan = plan
        self.dense = dense
        self.col_scale = col_scale
        self.diagonal = tuple(d.to(torch.float64).contiguous() for d in diagonal)
        p = plan
        stream = _lib.stream_handle(p.device)
        self.packed = []
        keep = []
        for k, (gi, slots, _) in enumerate(p.groups):
            grp = jacobian.groups[gi]
            g = p.plan.groups[k]
            if k == p.wgroup:
                blocks = [grp.blocks[0].contiguous(), grp.blocks[1].contiguous()]
            else:
                rs = [s for s, t in enumerate(grp.type_indices) if t != p.e]
                blocks = [grp.blocks[s].contiguous() for s in rs]
            keep.append(blocks)

            def pack(out, k=k, blocks=blocks):
                bufs = [None] * len(p.plan.groups)
                bufs[k] = out
                op = p.dev.operator(p.dtype, bufs, p.partials)
                ptrs = (ctypes.c_void_p * 3)(*[b.data_ptr() for b in blocks] + [None] * (3 - len(blocks)))
                if col_scale is None:
                    _lib.check(p.lib.tlh_pack_blocks(ctypes.byref(op), k, ptrs, _lib.dtype_code(blocks[0].dtype),
                                                     stream), "tlh_pack_blocks")
                else:
                    _lib.check(p.lib.tlh_pack_blocks_scaled(ctypes.byref(op), k, ptrs,
                                                            _lib.dtype_code(blocks[0].dtype), col_scale.data_ptr(),
                                                            stream), "tlh_pack_blocks_scaled")

            if k == p.wgroup:  # every slot of the jacobian group: shareable with other plans
                self.packed.append(packed_blocks(jacobian, gi, p.dev.order_keys[k], g.components, g.stride, p.dtype,
                                                 pack, col_scale))
            else:
                out = torch.empty((g.components, max(g.stride, 1)), dtype=p.dtype, device=p.device)
                pack(out)
                self.packed.append(out)
        self._keep = keep
        self.op = p.dev.operator(p.dtype, self.packed, p.partials)
        self.sop = p.sdev.operator(p.dtype, [self.packed[p.wgroup]], p.pre_partials)

    def damped(self, damping: float, *, defer_check: bool = False) -> "HipSchurSystem":
        return HipSchurSystem(self, damping, defer_check=defer_check)


class HipSchurSystem:
    """DampedSchurSystem (schur.py:259-397) on the HIP kernels.

    With ``defer_check`` the positive-definiteness flag of the eliminated-block
    inverses stays on the device; the preconditioner shares it and the HIP PCG
    reads it with its first scalar read, raising LinAlgError before any
    iteration (the reference raises from torch.linalg.cholesky at this point)."""

    hip_graph_safe = True  # matvec launches only HIP kernels on the current stream
    pcg_method = "matvec"  # described by pcg_operator for native PCG batches

    def __init__(self, lin: HipSchurLinearization, damping: float, *, defer_check: bool = False):
        self.lin, self.plan, self.damping = lin, lin.plan, float(damping)
        p = self.plan
        dev = p.device
        stream = _lib.stream_handle(dev)
        self.status = torch.zeros(1, dtype=torch.int32, device=dev)
        self.deferred = bool(defer_check)
        d = p.edim
        self.cinv = torch.empty(max(1, p.ecount * d * (d + 1) // 2), dtype=torch.float64, device=dev)
        _lib.check(p.lib.tlh_block_inverse(lin.diagonal[p.e].data_ptr(), p.ecount, d, self.damping,
                                           self.cinv.data_ptr(), self.status.data_ptr(), stream), "tlh_block_inverse")
        if not self.deferred:
            self.check()
        self._bp = None

    def check(self) -> None:
        if int(self.status.item()) != 0:
            raise LinAlgError("damped eliminated block is not positive definite")

    # -- helpers
    def _apply(self, mode, xc, bp, outp):
        p = self.plan
        _lib.check(p.lib.tlh_schur_apply(ctypes.byref(self.lin.op), p.wgroup, p.eslot, mode,
                                         None if xc is None else xc.data_ptr(), None if bp is None else bp.data_ptr(),
                                         None if outp is None else outp.data_ptr(), self.cinv.data_ptr(),
                                         _lib.stream_handle(p.device)), "tlh_schur_apply")

    def _others(self, x):
        p = self.plan
        for k in range(len(p.groups)):
            if k != p.wgroup:
                _lib.check(p.lib.tlh_group_tiles(ctypes.byref(self.lin.op), k, x.data_ptr(),
                                                 _lib.stream_handle(p.device)), "tlh_group_tiles")

    def _finish(self, base, scale):
        p = self.plan
        y = torch.empty(p.dim, dtype=p.dtype, device=p.device)
        _lib.check(p.lib.tlh_finish_types(ctypes.byref(self.lin.op), 0, p.ntypes_ret, base.data_ptr(), float(scale),
                                          y.data_ptr(), _lib.stream_handle(p.device)), "tlh_finish_types")
        return y

    # -- DampedSchurSystem interface
    def _copy(self, src, src_index, src_offset, dst, dst_index, dst_offset, count, scale=1.0):
        p = self.plan
        _lib.check(p.lib.tlh_copy_cast(src.data_ptr(), _lib.dtype_code(src.dtype),
                                       None if src_index is None else src_index.data_ptr(), src_offset,
                                       dst.data_ptr(), _lib.dtype_code(dst.dtype),
                                       None if dst_index is None else dst_index.data_ptr(), dst_offset, count,
                                       float(scale), _lib.stream_handle(p.device)), "tlh_copy_cast")

    def reduced_rhs(self, rhs: torch.Tensor, *, scale: float = 1.0) -> torch.Tensor:
        """Reduced right-hand side of ``scale * rhs`` (the LM passes the gradient and -1)."""
        p = self.plan
        if rhs.dtype not in (torch.float32, torch.float64):
            rhs = rhs.to(torch.float64)
        rhs = rhs.contiguous()
        self._bp = torch.empty(p.ecount * p.edim, dtype=p.dtype, device=p.device)
        self._copy(rhs, None, p.eoffset, self._bp, None, 0, self._bp.numel(), scale)
        b_ret = torch.empty(p.dim, dtype=p.dtype, device=p.device)
        self._copy(rhs, p.global_indices, 0, b_ret, None, 0, p.dim, scale)
        self._apply(_lib_mode("RHS"), None, self._bp, None)
        if len(p.groups) > 1:
            self._others(torch.zeros_like(b_ret))
        return self._finish(b_ret, 1.0)

    def matvec(self, vector: torch.Tensor) -> torch.Tensor:
        p = self.plan
        if vector.shape != (p.dim,) or vector.dtype != p.dtype:
            raise ValueError("Schur matvec input must be the retained vector in the plan dtype")
        x = vector.contiguous()
        self._apply(_lib_mode("MATVEC"), x, None, None)
        self._others(x)
        return self._finish(x, self.damping)

    def pcg_operator(self, damping=None) -> "_lib.PcgOperator":
        """Descriptor of ``matvec`` for the native PCG batch (tlh_pcg_iterations); the
        system's own damping applies (``damping`` is ignored)."""
        p = self.plan
        A = _lib.PcgOperator()
        A.kind = _lib.PCG_OP_SCHUR
        A.wgroup, A.eslot, A.type_end = p.wgroup, p.eslot, p.ntypes_ret
        A.op = ctypes.pointer(self.lin.op)
        A.cinv = self.cinv.data_ptr()
        A.damping = self.damping
        return A

    def backsubstitute(self, retained: torch.Tensor) -> torch.Tensor:
        p = self.plan
        if self._bp is None:
            raise ValueError("reduced_rhs must be called before backsubstitute")
        xr = retained.to(p.dtype).contiguous()
        xe = torch.empty(p.ecount * p.edim, dtype=p.dtype, device=p.device)
        self._apply(_lib_mode("BACKSUB"), xr, self._bp, xe)
        if p.lonely is not None:  # the tile pass writes only the observed eliminated variables
            _lib.check(p.lib.tlh_block_apply(self.cinv.data_ptr(), p.edim, p.lonely.data_ptr(), p.lonely.numel(),
                                             self._bp.data_ptr(), xe.data_ptr(), _lib.dtype_code(p.dtype),
                                             _lib.stream_handle(p.device)), "tlh_block_apply")
        # the retained indices and the eliminated range cover the whole tangent vector
        total = sum(e.count * e.dim for e in p.layout)
        out = torch.empty(total, dtype=torch.float64, device=p.device)
        self._copy(xr, None, 0, out, p.global_indices, 0, p.dim)
        self._copy(xe, None, 0, out, None, p.eoffset, xe.numel())
        return out

    def preconditioner_blocks(self):
        """float64 retained blocks D + lambda I - sum W_e C^-1 W_e^T (schur.py:353-362)."""
        p = self.plan
        blocks = []
        w = p.sparse_plan.groups[0]
        wtype = w.slots[1 - p.eslot].type
        lds = _lib.device_info(p.device).lds_per_block
        for k, t in enumerate(p.retained):
            D = self.lin.diagonal[t]
            # the camera-order kernel reads the raw (unscaled) blocks: a scaled system uses the tile kernel
            if k == wtype and p.major is not None and self.lin.col_scale is None:
                M = torch.empty_like(D)
                order, edge_ptr, var_edge_ptr, edge_point = p.major
                raw = self.lin._keep[p.wgroup]
                jc, jp = raw[1 - p.eslot], raw[p.eslot]
                _lib.check(p.lib.tlh_schur_precond_major(
                    jc.data_ptr(), jp.data_ptr(), _lib.dtype_code(jc.dtype), 2, 6, 3, order.data_ptr(),
                    edge_ptr.data_ptr(), var_edge_ptr.data_ptr(), edge_point.data_ptr(), D.shape[0],
                    self.cinv.data_ptr(), D.data_ptr(), self.damping, M.data_ptr(), _lib.stream_handle(p.device)),
                    "tlh_schur_precond_major")
            elif k == wtype:
                M = torch.empty_like(D)
                _lib.check(p.lib.tlh_schur_precond(ctypes.byref(self.lin.sop), 0, p.eslot, self.cinv.data_ptr(),
                                                   D.data_ptr(), self.damping, M.data_ptr(), lds,
                                                   _lib.stream_handle(p.device)), "tlh_schur_precond")
            else:
                M = D + self.damping * torch.eye(D.shape[-1], dtype=D.dtype, device=D.device)
            blocks.append(M)
        return blocks

    def preconditioner(self, *, graph_safe=False):
        from .hip_diagonal import HipBlockJacobi
        from .sparse import VariableLayout

        p = self.plan
        lay = tuple(VariableLayout(p.layout[t].var_type, p.layout[t].ids, off)
                    for t, (_, _, off) in zip(p.retained, p.rlayout[:-1]))
        return HipBlockJacobi(lay, self.preconditioner_blocks(), 0.0, status=self.status, defer_check=self.deferred)

    def dense_matrix(self) -> torch.Tensor:
        p = self.plan
        pairs = p._dense_pairs()
        n = p.dim
        M = torch.zeros((n, n), dtype=torch.float64, device=p.device)
        for t, (c, d, off) in zip(p.retained, p.rlayout[:-1]):
            # block v of this type sits at rows and columns off + v*d .. off + v*d + d - 1
            blocks = M.as_strided((c, d, d), (d * n + d, n, 1), storage_offset=off * n + off)
            blocks.copy_(self.lin.diagonal[t])
            blocks.diagonal(dim1=1, dim2=2).add_(self.damping)
        wg = p.plan.groups[p.wgroup]
        # per-observation W_o and W_o C_p^-1 (float64), read by the camera-pair sums
        work = torch.empty(max(1, wg.count * 2 * wg.slots[1 - p.eslot].dim * p.edim), dtype=torch.float64,
                           device=p.device)
        _lib.check(p.lib.tlh_schur_dense(ctypes.byref(self.lin.op), p.wgroup, p.eslot, self.cinv.data_ptr(),
                                         pairs["ptr"].data_ptr(), pairs["i"].data_ptr(), pairs["j"].data_ptr(),
                                         pairs["ci"].data_ptr(), pairs["cj"].data_ptr(), pairs["n"], M.data_ptr(), n,
                                         work.data_ptr(), _lib.stream_handle(p.device)), "tlh_schur_dense")
        return M

    def dense_solve(self, rhs: torch.Tensor) -> torch.Tensor:
        from .cholesky import cholesky_solve_f64

        if self.deferred:
            self.check()
        return cholesky_solve_f64(self.dense_matrix(), rhs.to(torch.float64)).to(self.plan.dtype)


def _lib_mode(name: str) -> int:
    return {"MATVEC": 0, "RHS": 1, "BACKSUB": 2}[name]
