"""Symbolic plan of the deterministic tile reduction (host side, numpy).

Built once per solve from the incidence structure only -- the variable
indices of every slot -- mirroring the reference's split between a symbolic
plan and per-linearization values (warp_normal.py:5-6, 125-185;
warp_diagonal.py:3-4, 70-121).  Where the reference shards "hot" columns over
128 atomic accumulators (warp_normal.py:168-176), this plan instead fixes a
summation order for every variable:

* each residual group is sorted by its *primary* slot (the slot whose variable
  type has the most variables, e.g. points in bundle adjustment), then by the
  remaining slots (so repeated (camera, point) edges become adjacent), then by
  the original position (stable);
* the sorted observations are cut into tiles of at most ``tile_cap``
  observations, at primary-variable boundaries whenever the run fits;
* for every slot, each tile is ordered by that slot's variable (``rank``) and
  split into segments (runs of one variable, at most ``piece`` long for
  SPARSE slots);
* a segment's sum goes to a *partial* (SPARSE) or into a per-workgroup dense
  accumulator of the whole type (DENSE, when the type fits in LDS);
* every variable's partials are finally summed in increasing partial id order
  (dense blocks first, then its CSR list).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

REDUCE_SPARSE = 0
REDUCE_DENSE = 1
THREADS = 256


def items_for(esize: int) -> int:
    """Observations per thread per tile of the specialised kernels (csrc/normal.hip)."""
    return 4 if esize == 4 else 2


WAVE_THREADS = 512  # TLH_WAVE_THREADS
WAVE_TILE = 1024  # TLH_WAVE_TILE (float32; half for float64)


def wave_items(esize: int) -> int:
    """Consecutive positions per thread of the wave kernel (csrc/normal.hip)."""
    return WAVE_TILE // WAVE_THREADS // (esize // 4)


@dataclass
class SlotPlan:
    type: int
    dim: int
    mode: int
    index: np.ndarray  # int32 [N] sorted order
    rank: np.ndarray | None  # uint16 [N] or None (identity)
    seg_ptr: np.ndarray  # int32 [ntiles+1]
    seg_start: np.ndarray  # int32 [nseg] tile-local
    seg_var: np.ndarray  # int64 [nseg] variable of each segment
    seg_target: np.ndarray | None = None  # int32 [nseg] partial id / variable (-1: direct write)
    dense_block: int = 0
    pos_base: int = 0  # wave groups: first positional partial id of this slot
    # wave groups, DENSE secondary slot reduced in occurrence rounds: rank then holds
    # each position's occurrence index (earlier positions of the tile with the same
    # variable) and rounds[t] = 1 + the largest occurrence index of tile t
    rounds: np.ndarray | None = None  # int32 [ntiles]


@dataclass
class GroupPlan:
    count: int
    rows: int
    primary: int
    order: np.ndarray  # int64 [N] original observation at each sorted position
    tile_start: np.ndarray  # int32 [ntiles+1]
    tile_cap: int
    nwg: int
    tiles_per_wg: int
    slots: list
    kernel: int = 0  # 0: segment-array kernel, 1: wave kernel (2 slots, specialised shape)

    @property
    def ntiles(self) -> int:
        return int(self.tile_start.size - 1)

    @property
    def stride(self) -> int:
        # rounded up to a whole wave-kernel tile: its loads never need bounds branches
        return int(-(-self.count // WAVE_TILE) * WAVE_TILE) if self.count else 0

    @property
    def components(self) -> int:
        return self.rows * sum(s.dim for s in self.slots)


@dataclass
class TypePlan:
    count: int
    dim: int
    offset: int
    dense_blocks: int = 0
    sparse_partials: int = 0
    csr_ptr: np.ndarray = field(default_factory=lambda: np.zeros(1, np.int32))
    csr_idx: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    direct: np.ndarray | None = None  # uint8 [count]: written by its only wave piece
    finish_list: np.ndarray | None = None  # int32 variables the finish kernel writes (None: all)

    @property
    def partial_count(self) -> int:
        return self.dense_blocks * self.count + self.sparse_partials


@dataclass
class StructurePlan:
    types: list
    groups: list
    esize: int
    tile_cap: int

    @property
    def size(self) -> int:
        return sum(t.count * t.dim for t in self.types)


def _cut_tiles(primary_sorted: np.ndarray, cap: int) -> np.ndarray:
    n = primary_sorted.size
    if n == 0:
        return np.zeros(1, dtype=np.int32)
    run_starts = np.concatenate(([0], np.flatnonzero(np.diff(primary_sorted)) + 1))
    starts = [0]
    s = 0
    while s < n:
        target = s + cap
        if target >= n:
            end = n
        else:
            j = int(np.searchsorted(run_starts, target, side="right")) - 1
            end = int(run_starts[j])
            if end <= s:  # one variable's run is longer than a tile: split it
                end = target
        starts.append(end)
        s = end
    return np.asarray(starts, dtype=np.int64)


def _slot_segments(index_sorted: np.ndarray, tile_start: np.ndarray, piece: int | None):
    """Within-tile order by variable; rank, segment starts/tiles/variables."""
    n = index_sorted.size
    ntiles = tile_start.size - 1
    lens = np.diff(tile_start)
    tile_of = np.repeat(np.arange(ntiles, dtype=np.int64), lens)
    pos = np.arange(n, dtype=np.int64)
    perm = np.lexsort((pos, index_sorted, tile_of))
    local = pos - tile_start[tile_of[perm]]  # position k in tile order (perm index k)
    rank = np.empty(n, dtype=np.int64)
    rank[perm] = local
    v = index_sorted[perm]
    t = tile_of[perm]
    boundary = np.ones(n, dtype=bool)
    boundary[1:] = (v[1:] != v[:-1]) | (t[1:] != t[:-1])
    if piece is not None and n:
        run_id = np.cumsum(boundary) - 1
        run_begin = np.flatnonzero(boundary)
        in_run = pos - run_begin[run_id]
        boundary |= (in_run % piece) == 0
    heads = np.flatnonzero(boundary)
    seg_tile = t[heads]
    seg_ptr = np.zeros(ntiles + 1, dtype=np.int64)
    np.cumsum(np.bincount(seg_tile, minlength=ntiles), out=seg_ptr[1:])
    identity = bool(np.all(rank == (pos - tile_start[tile_of])))
    return rank, identity, local[heads], seg_ptr, v[heads]


WAVE = 64


def _aligned_tiles(n: int, cap: int) -> np.ndarray:
    """Tiles of exactly ``cap`` sorted positions (the last one may be short)."""
    return np.append(np.arange(0, n, cap, dtype=np.int64), n) if n else np.zeros(1, dtype=np.int64)


# Most rounds a tile may take in the occurrence-round reduction; groups with a
# more repetitive secondary slot (few variables) use the staged reduction.
MAX_ROUNDS = 8


def _occurrences(index_sorted: np.ndarray, tile_start: np.ndarray):
    """Occurrence index of each position's variable inside its tile (0 for the
    first position of the tile with that variable, in position order) and the
    number of rounds (1 + largest occurrence index) of each tile."""
    n = index_sorted.size
    ntiles = tile_start.size - 1
    if n == 0:
        return np.zeros(0, np.int64), np.zeros(ntiles, np.int32)
    tile = np.repeat(np.arange(ntiles, dtype=np.int64), np.diff(tile_start))
    key = tile * (int(index_sorted.max()) + 1) + index_sorted.astype(np.int64)
    order = np.argsort(key, kind="stable")
    ks = key[order]
    first = np.r_[True, ks[1:] != ks[:-1]]
    start = np.maximum.accumulate(np.where(first, np.arange(n), 0))
    occ = np.empty(n, np.int64)
    occ[order] = np.arange(n) - start
    rounds = np.zeros(ntiles, np.int64)
    np.maximum.at(rounds, tile, occ + 1)
    return occ, rounds.astype(np.int32)


def _wave_pieces(index_sorted: np.ndarray, tile_start: np.ndarray, span: int):
    """Primary-slot pieces of the wave kernel: runs of one variable, also cut at
    every ``span``-position boundary of each tile (one wavefront covers 64
    threads x ITEMS consecutive positions)."""
    n = index_sorted.size
    ntiles = tile_start.size - 1
    tile_of = np.repeat(np.arange(ntiles, dtype=np.int64), np.diff(tile_start))
    pos = np.arange(n, dtype=np.int64)
    local = pos - tile_start[tile_of]
    head = np.ones(n, dtype=bool)
    head[1:] = index_sorted[1:] != index_sorted[:-1]
    head |= (local % span) == 0
    heads = np.flatnonzero(head)
    ends = np.append(heads[1:], n) - 1
    seg_ptr = np.zeros(ntiles + 1, dtype=np.int64)
    np.cumsum(np.bincount(tile_of[heads], minlength=ntiles), out=seg_ptr[1:])
    return local[heads], seg_ptr, index_sorted[heads], ends


def build_structure(
    layout: list,
    groups: list,
    *,
    esize: int,
    lds_limit: int = 65536,
    cu_count: int = 256,
    wgs_per_cu: int = 2,
    piece: int = 32,
    tile_cap: int | None = None,
    allow_dense: bool = True,
    force_primary: dict | None = None,
    split_primary: bool = True,
    wave_shapes: frozenset = frozenset(),
) -> StructurePlan:
    """Plan the deterministic reduction.

    ``layout``: [(count, dim, offset)] per variable type.
    ``groups``: [(rows, [type per slot], [int64 index array per slot])].
    """
    cap = tile_cap or THREADS * items_for(esize)
    types = [TypePlan(int(c), int(d), int(o)) for c, d, o in layout]
    gplans = []
    for gi, (rows, slot_types, slot_idx) in enumerate(groups):
        ns = len(slot_types)
        if not 1 <= ns <= 3:
            raise ValueError("the hip backend supports 1 to 3 variable slots per residual")
        count = int(slot_idx[0].size)
        sizes = [types[t].count for t in slot_types]
        primary = int(np.argmax(sizes))
        if force_primary and gi in force_primary:
            primary = int(force_primary[gi])
        keys = [np.arange(count)]
        keys += [slot_idx[s] for s in reversed(range(ns)) if s != primary]
        keys.append(slot_idx[primary])
        order = np.lexsort(keys).astype(np.int64) if count else np.zeros(0, np.int64)
        sorted_idx = [np.asarray(slot_idx[s])[order] for s in range(ns)]
        dmax = max(types[t].dim for t in slot_types)
        shape = (int(rows), *(types[t].dim for t in slot_types))
        wave = ns == 2 and shape in wave_shapes and count > 0
        gcap = WAVE_THREADS * wave_items(esize) if wave else cap
        tile_start = _aligned_tiles(count, gcap) if wave else _cut_tiles(sorted_idx[primary], cap)
        ntiles = tile_start.size - 1
        stage = gcap * dmax * esize + (gcap * 4 if wave else 0)
        dense_budget = lds_limit - stage
        slots = []
        ends = None
        for s in range(ns):
            t = slot_types[s]
            dense_bytes = types[t].count * types[t].dim * esize
            dense = allow_dense and s != primary and ntiles > 1 and dense_bytes <= dense_budget
            if dense:
                dense_budget -= dense_bytes
            split = None if dense or (s == primary and not split_primary) or wave else piece
            if wave and s == primary:
                seg_start, seg_ptr, seg_var, ends = _wave_pieces(sorted_idx[s], tile_start,
                                                                 WAVE * wave_items(esize))
                rank, identity = None, True
            else:
                rank, identity, seg_start, seg_ptr, seg_var = _slot_segments(sorted_idx[s], tile_start, split)
            slots.append(
                SlotPlan(
                    t,
                    types[t].dim,
                    REDUCE_DENSE if dense else REDUCE_SPARSE,
                    sorted_idx[s].astype(np.int32),
                    None if identity else rank.astype(np.uint16),
                    seg_ptr.astype(np.int32),
                    seg_start.astype(np.int32),
                    seg_var.astype(np.int64),
                )
            )
            if wave and s == primary:
                slots[-1].piece_end = ends
            elif wave and dense:
                occ, rounds = _occurrences(sorted_idx[s], tile_start)
                if rounds.size and int(rounds.max()) <= MAX_ROUNDS:
                    slots[-1].rank = occ.astype(np.uint16)
                    slots[-1].rounds = rounds
        has_dense = any(sl.mode == REDUCE_DENSE for sl in slots)
        if has_dense and ntiles:
            # a wave-kernel workgroup (1024 threads) fills its compute unit's wave slots
            nwg = min(ntiles, cu_count * (1 if wave else wgs_per_cu))
            per = -(-ntiles // nwg)
            nwg = -(-ntiles // per)
        else:
            nwg, per = ntiles, 1
        gplans.append(
            GroupPlan(count, int(rows), primary, order, tile_start.astype(np.int32), gcap, nwg, per, slots,
                      kernel=1 if wave else 0)
        )
    # dense blocks first (ids b*count + v), then sparse partials
    for g in gplans:
        for sl in g.slots:
            if sl.mode == REDUCE_DENSE:
                sl.dense_block = types[sl.type].dense_blocks
                types[sl.type].dense_blocks += g.nwg
                sl.seg_target = sl.seg_var.astype(np.int32)
    pairs = [[] for _ in types]
    next_id = [t.dense_blocks * t.count for t in types]
    for g in gplans:
        for sl in g.slots:
            if sl.mode == REDUCE_SPARSE and g.kernel == 0:
                ids = next_id[sl.type] + np.arange(sl.seg_var.size, dtype=np.int64)
                next_id[sl.type] += sl.seg_var.size
                sl.seg_target = ids.astype(np.int32)
                pairs[sl.type].append((sl.seg_var, ids))
    # wave groups: positional partial ids (one region of `count` ids per slot)
    for g in gplans:
        if g.kernel != 1:
            continue
        for si, sl in enumerate(g.slots):
            if sl.mode == REDUCE_DENSE:
                continue
            sl.pos_base = next_id[sl.type]
            next_id[sl.type] += g.count
            if si == g.primary:
                ids = sl.pos_base + sl.piece_end
            else:
                tile_of = np.repeat(np.arange(g.ntiles, dtype=np.int64), np.diff(sl.seg_ptr))
                ids = sl.pos_base + g.tile_start[tile_of].astype(np.int64) + sl.seg_start
            sl.seg_target = ids.astype(np.int64)
            pairs[sl.type].append((sl.seg_var, ids))
    # a variable whose only contribution is one primary wave piece is written directly
    for t, ty in enumerate(types):
        contrib = np.zeros(ty.count, dtype=np.int64)
        if ty.dense_blocks:
            contrib += 2
        for var, _ in pairs[t]:
            contrib += np.bincount(var, minlength=ty.count)
        direct = np.zeros(ty.count, dtype=bool)
        for g in gplans:
            if g.kernel == 1:
                sl = g.slots[g.primary]
                if sl.type == t:
                    cand = np.zeros(ty.count, dtype=bool)
                    cand[sl.seg_var] = True
                    direct |= cand & (contrib == 1)
        if direct.any():
            ty.direct = direct.astype(np.uint8)
            ty.finish_list = np.flatnonzero(~direct).astype(np.int32)
            keep = []
            for var, ids in pairs[t]:
                m = ~direct[var]
                keep.append((var[m], ids[m]))
            pairs[t] = keep
            for g in gplans:
                if g.kernel == 1 and g.slots[g.primary].type == t:
                    sl = g.slots[g.primary]
                    sl.seg_target = np.where(direct[sl.seg_var], -1, sl.seg_target)
    for g in gplans:
        for sl in g.slots:
            if sl.seg_target is not None and sl.seg_target.dtype != np.int32:
                sl.seg_target = sl.seg_target.astype(np.int32)
    for t, ty in enumerate(types):
        ty.sparse_partials = next_id[t] - ty.dense_blocks * ty.count
        if next_id[t] > 2**31 - 1:
            raise ValueError("partial ids exceed the 32-bit range")
        if pairs[t]:
            var = np.concatenate([p[0] for p in pairs[t]])
            ids = np.concatenate([p[1] for p in pairs[t]])
            o = np.lexsort((ids, var))
            ty.csr_idx = ids[o].astype(np.int32)
            ptr = np.zeros(ty.count + 1, dtype=np.int64)
            np.cumsum(np.bincount(var, minlength=ty.count), out=ptr[1:])
            ty.csr_ptr = ptr.astype(np.int32)
        else:
            ty.csr_ptr = np.zeros(ty.count + 1, dtype=np.int32)
    return StructurePlan(types, gplans, esize, cap)


# Variable-major accumulation (csrc/major.hip) pays off for a type with few
# variables observed many times each: its blocks then need no per-tile partials.
MAJOR_MAX_VARIABLES = 65536
MAJOR_MIN_MEAN = 64


def major_eligible(variables: int, observations: int) -> bool:
    return 0 < variables <= MAJOR_MAX_VARIABLES and observations >= MAJOR_MIN_MEAN * variables \
        and observations < 2**31


def major_order(index: np.ndarray, count: int):
    """Observations sorted by variable (stable) and each variable's range: (order, ptr), int32."""
    index = np.asarray(index, dtype=np.int64).reshape(-1)
    order = np.argsort(index, kind="stable")
    ptr = np.zeros(count + 1, dtype=np.int64)
    np.cumsum(np.bincount(index, minlength=count), out=ptr[1:])
    return order.astype(np.int32), ptr.astype(np.int32)


def major_edges(index_c: np.ndarray, index_p: np.ndarray, count_c: int):
    """Observations sorted by (camera, point, position) and their (camera, point) runs:
    (order, edge_ptr [E+1], var_edge_ptr [count_c+1], edge_point [E]), all int32."""
    ic = np.asarray(index_c, dtype=np.int64).reshape(-1)
    ip = np.asarray(index_p, dtype=np.int64).reshape(-1)
    n = ic.size
    order = np.lexsort((np.arange(n), ip, ic))
    sc, sp = ic[order], ip[order]
    head = np.ones(n, dtype=bool)
    if n:
        head[1:] = (sc[1:] != sc[:-1]) | (sp[1:] != sp[:-1])
    starts = np.flatnonzero(head)
    edge_ptr = np.append(starts, n)
    edge_cam = sc[starts]
    var_edge_ptr = np.zeros(count_c + 1, dtype=np.int64)
    np.cumsum(np.bincount(edge_cam, minlength=count_c), out=var_edge_ptr[1:])
    return (order.astype(np.int32), edge_ptr.astype(np.int32), var_edge_ptr.astype(np.int32),
            sp[starts].astype(np.int32))


def emulate_normal_matvec(plan: StructurePlan, packed: list, x: np.ndarray, damping: float):
    """Execute the plan's data flow on the CPU (numpy), exactly as the kernels do.

    ``packed[g]`` is the group's component-major array [components, count] in
    sorted order.  Used by the CPU tests to validate plans; not a product path.
    """
    types = plan.types
    parts = [np.zeros((t.partial_count, t.dim)) for t in types]
    direct_out = []
    for g, gp in zip(plan.groups, packed):
        offs = np.cumsum([0] + [s.dim for s in g.slots])
        R = g.rows
        xs = []
        for s in g.slots:
            t = types[s.type]
            xs.append(x[t.offset : t.offset + t.count * t.dim].reshape(t.count, t.dim)[s.index])
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
            for tile in range(g.ntiles):
                a, b = g.tile_start[tile], g.tile_start[tile + 1]
                stage = np.zeros((b - a, s.dim))
                rank = np.arange(b - a) if s.rank is None else s.rank[a:b].astype(np.int64)
                stage[rank] = cont[a:b]
                k0, k1 = s.seg_ptr[tile], s.seg_ptr[tile + 1]
                for k in range(k0, k1):
                    lo = s.seg_start[k]
                    hi = s.seg_start[k + 1] if k + 1 < k1 else b - a
                    total = stage[lo:hi].sum(axis=0)
                    if s.mode == REDUCE_DENSE:
                        dense[si][tile // g.tiles_per_wg, s.seg_target[k]] += total
                    elif s.seg_target[k] < 0:
                        v = s.seg_var[k]
                        direct_out.append((t, v, total))
                    else:
                        parts[s.type][s.seg_target[k]] = total
            if s.mode == REDUCE_DENSE:
                for w in range(g.nwg):
                    base = (s.dense_block + w) * t.count
                    parts[s.type][base : base + t.count] = dense[si][w]
    y = np.zeros_like(x)
    for ti, t in enumerate(types):
        acc = np.zeros((t.count, t.dim))
        for b in range(t.dense_blocks):
            acc += parts[ti][b * t.count : (b + 1) * t.count]
        for v in range(t.count):
            for k in range(t.csr_ptr[v], t.csr_ptr[v + 1]):
                acc[v] += parts[ti][t.csr_idx[k]]
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
        cols = [np.asarray(b)[g.order].reshape(g.count, -1).T for b in blocks]
        out.append(np.concatenate(cols, axis=0) if cols else np.zeros((0, g.count)))
    return out
