"""float64 Schur elimination of one independent variable type (CPU oracle).

Restates torchls.schur (schur.py:17-397) without CSR storage: with A = J^T J,
eliminated type e (e.g. points) and retained variables c,

    C   = D_e + lambda I                (per eliminated variable, schur.py:266-269)
    W y = sum_obs J_c^T (sum_e J_e) y   (cross blocks, schur.py:213-224)
    S x = A_cc x + lambda x - W C^-1 W^T x               (schur.py:303-328)
    rhs_c' = rhs_c - W C^-1 rhs_e                        (schur.py:285-290)
    x_e = C^-1 (rhs_e - W^T x_c)                         (schur.py:292-301)
    M_c = D_c + lambda I - sum_e W_ce C_e^-1 W_ce^T      (schur.py:353-362)

Pinned by the dense closed form S = A_cc + lambda I - A_ce (A_ee + lambda I)^-1 A_ec.
"""

from __future__ import annotations

import numpy as np

from .sparse import NpJacobian


class SchurSystem:
    def __init__(self, J: NpJacobian, diagonal: list, eliminated: int, damping: float):
        self.J, self.e, self.lam = J, eliminated, damping
        ce, de, oe = J.layout[eliminated]
        self.ce, self.de, self.oe = ce, de, oe
        self.retained = [t for t in range(len(J.layout)) if t != eliminated]
        # retained layout: concatenated in type order
        self.rlayout, off = [], 0
        for t in self.retained:
            c, d, _ = J.layout[t]
            self.rlayout.append((c, d, off))
            off += c * d
        self.dim = off
        C = diagonal[eliminated] + damping * np.eye(de)
        self.Cinv = np.linalg.inv(C)
        self.diag = diagonal

    def _split(self, v_full):
        return np.concatenate([v_full[o : o + c * d] for t in self.retained for c, d, o in [self.J.layout[t]]])

    def _ret_col(self, t):
        """Offset of type t inside the retained vector."""
        return self.rlayout[self.retained.index(t)][2]

    def _group_parts(self, g):
        es = [s for s, t in enumerate(g.types) if t == self.e]
        rs = [s for s, t in enumerate(g.types) if t != self.e]
        return es, rs

    def Wt(self, x_ret):
        """W^T x for retained x: [ce, de]."""
        out = np.zeros((self.ce, self.de))
        for g in self.J.groups:
            es, rs = self._group_parts(g)
            if not es or not rs:
                continue
            E = sum(g.blocks[s] for s in es)
            q = np.zeros((g.count, g.rows))
            for s in rs:
                t = g.types[s]
                d = self.J.layout[t][1]
                cols = self._ret_col(t) + g.indices[s][:, None] * d + np.arange(d)
                q += np.einsum("nrd,nd->nr", g.blocks[s], x_ret[cols])
            contrib = np.einsum("nre,nr->ne", E, q)
            np.add.at(out, g.indices[es[0]], contrib)
        return out

    def W(self, y):
        """W y for y [ce, de] -> retained vector."""
        out = np.zeros(self.dim)
        for g in self.J.groups:
            es, rs = self._group_parts(g)
            if not es or not rs:
                continue
            E = sum(g.blocks[s] for s in es)
            q = np.einsum("nre,ne->nr", E, y[g.indices[es[0]]])
            for s in rs:
                t = g.types[s]
                d = self.J.layout[t][1]
                cols = self._ret_col(t) + g.indices[s][:, None] * d + np.arange(d)
                np.add.at(out, cols.ravel(), np.einsum("nrd,nr->nd", g.blocks[s], q).ravel())
        return out

    def B(self, x_ret):
        """(J_ret^T J_ret) x over retained slots of every group."""
        out = np.zeros(self.dim)
        for g in self.J.groups:
            rs = [s for s, t in enumerate(g.types) if t != self.e]
            if not rs:
                continue
            q = np.zeros((g.count, g.rows))
            colss = []
            for s in rs:
                t = g.types[s]
                d = self.J.layout[t][1]
                cols = self._ret_col(t) + g.indices[s][:, None] * d + np.arange(d)
                colss.append(cols)
                q += np.einsum("nrd,nd->nr", g.blocks[s], x_ret[cols])
            for s, cols in zip(rs, colss):
                np.add.at(out, cols.ravel(), np.einsum("nrd,nr->nd", g.blocks[s], q).ravel())
        return out

    def matvec(self, x_ret):
        y = np.einsum("nij,nj->ni", self.Cinv, self.Wt(x_ret))
        return self.B(x_ret) + self.lam * x_ret - self.W(y)

    def reduced_rhs(self, rhs_full):
        be = rhs_full[self.oe : self.oe + self.ce * self.de].reshape(self.ce, self.de)
        self._be = be
        return self._split(rhs_full) - self.W(np.einsum("nij,nj->ni", self.Cinv, be))

    def backsubstitute(self, x_ret):
        xe = np.einsum("nij,nj->ni", self.Cinv, self._be - self.Wt(x_ret))
        out = np.zeros(self.J.n)
        k = 0
        for t in self.retained:
            c, d, o = self.J.layout[t]
            out[o : o + c * d] = x_ret[k : k + c * d]
            k += c * d
        out[self.oe : self.oe + self.ce * self.de] = xe.reshape(-1)
        return out

    def preconditioner_blocks(self):
        """Per retained type: D + lambda I - sum over unique (retained, e) edges W_ce C^-1 W_ce^T."""
        blocks = []
        for t in self.retained:
            c, d, _ = self.J.layout[t]
            Mt = self.diag[t] + self.lam * np.eye(d)
            edges = {}
            for g in self.J.groups:
                es, rs = self._group_parts(g)
                if not es:
                    continue
                E = sum(g.blocks[s] for s in es)
                for s in rs:
                    if g.types[s] != t:
                        continue
                    w = np.einsum("nri,nrj->nij", g.blocks[s], E)
                    keys = g.indices[s] * self.ce + g.indices[es[0]]
                    for k in range(g.count):
                        edges[int(keys[k])] = edges.get(int(keys[k]), 0) + w[k]
            for key, w in edges.items():
                ci, ei = divmod(key, self.ce)
                Mt[ci] -= w @ self.Cinv[ei] @ w.T
            blocks.append(Mt)
        return blocks

    def dense_matrix(self):
        n = self.dim
        S = np.zeros((n, n))
        for j in range(n):
            e = np.zeros(n)
            e[j] = 1.0
            S[:, j] = self.matvec(e)
        return 0.5 * (S + S.T)
