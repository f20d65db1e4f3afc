"""Symbolic plan of the deterministic tile reduction, built on the device.

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

Every array is a torch tensor on the device of the index arrays it is built
from, and every O(N) step runs there as a device primitive (stable radix
sorts, scans, histograms, gathers and scatters) -- the reference's plans are
built the same way from torch ops on the device (warp_normal.py:155-176,
schur.py:83-136).  The host reads back only scalars (sizes, maxima, flags) and,
for groups whose tiles are cut at primary-variable runs, the run starts that
the greedy tile cut walks.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

REDUCE_SPARSE = 0
REDUCE_DENSE = 1
THREADS = 256


def items_for(esize: int) -> int:
    """Observations per thread per tile of the specialised kernels (csrc/normal.hip)."""
    return 4 if esize == 4 else 2


WAVE_THREADS = 512  # TLH_WAVE_THREADS
WAVE_TILE = 2048  # TLH_WAVE_TILE (float32; half for float64)


def wave_items(esize: int) -> int:
    """Consecutive positions per thread of the wave kernel (csrc/normal.hip)."""
    return WAVE_TILE // WAVE_THREADS // (esize // 4)


@dataclass
class SlotPlan:
    type: int
    dim: int
    mode: int
    index: torch.Tensor  # int32 [N] sorted order
    rank: torch.Tensor | None  # int32 [N] (values < 65,536) or None (identity)
    seg_ptr: torch.Tensor  # int32 [ntiles+1]
    seg_start: torch.Tensor  # int32 [nseg] tile-local
    seg_var: torch.Tensor  # int64 [nseg] variable of each segment
    seg_target: torch.Tensor | None = None  # int32 [nseg] partial id / variable (-1: direct write)
    dense_block: int = 0
    pos_base: int = 0  # wave groups: first positional partial id of this slot
    # wave groups, DENSE secondary slot reduced in occurrence rounds: rank then holds
    # each position's occurrence index (earlier positions of the tile with the same
    # variable) and rounds[t] = 1 + the largest occurrence index of tile t
    rounds: torch.Tensor | None = None  # int32 [ntiles]


@dataclass
class GroupPlan:
    count: int
    rows: int
    primary: int
    order: torch.Tensor  # int64 [N] original observation at each sorted position
    tile_start: torch.Tensor  # int32 [ntiles+1]
    tile_cap: int
    nwg: int
    tiles_per_wg: int
    slots: list
    kernel: int = 0  # 0: segment-array kernel, 1: wave kernel (2 slots, specialised shape)
    # wave kernel, overflow encoding of the dense secondary slot (0: none): most
    # primary-slot piece ends (y / partial writes) of one wavefront over the tiles
    # of one workgroup, the capacity of its LDS output region, and most later
    # occurrences of one tile, the capacity of the overflow stage
    out_entries: int = 0
    stage_entries: int = 0

    @property
    def ntiles(self) -> int:
        return int(self.tile_start.numel() - 1)

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
    csr_ptr: torch.Tensor | None = None  # int32 [count+1]
    csr_idx: torch.Tensor | None = None  # int32 partial ids
    direct: torch.Tensor | None = None  # uint8 [count]: written by its only wave piece
    finish_list: torch.Tensor | None = None  # int32 variables the finish kernel writes (None: all)

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


def _arange(n, device):
    return torch.arange(n, dtype=torch.int64, device=device)


def _flatnonzero(mask):
    return torch.nonzero(mask).reshape(-1)


def _counts_ptr(values, count, device):
    """int64 [count+1] prefix sums of the histogram of ``values`` (CSR row pointers)."""
    ptr = torch.zeros(count + 1, dtype=torch.int64, device=device)
    if values.numel():
        ptr[1:] = torch.cumsum(torch.bincount(values, minlength=count), 0)
    return ptr


def _stable_order(keys):
    """Positions sorted lexicographically by ``keys`` (most significant first), ties in
    position order: successive stable sorts from the least significant key."""
    order = _arange(keys[0].numel(), keys[0].device)
    for k in reversed(keys):
        order = order[torch.sort(k[order], stable=True).indices]
    return order


def _cut_tiles(primary_sorted: torch.Tensor, cap: int) -> torch.Tensor:
    """Tiles of at most ``cap`` positions cut at primary-variable run starts where the run
    fits (greedy, one tile after the other): the run starts are read back and walked on
    the host."""
    n = primary_sorted.numel()
    device = primary_sorted.device
    if n == 0:
        return torch.zeros(1, dtype=torch.int64, device=device)
    head = torch.ones(n, dtype=torch.bool, device=device)
    head[1:] = primary_sorted[1:] != primary_sorted[:-1]
    run_starts = _flatnonzero(head).cpu().numpy()
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
    return torch.tensor(starts, dtype=torch.int64, device=device)


def _tile_of(tile_start: torch.Tensor, n: int) -> torch.Tensor:
    ntiles = tile_start.numel() - 1
    return torch.repeat_interleave(_arange(ntiles, tile_start.device), torch.diff(tile_start), output_size=n)


def _slot_segments(index_sorted: torch.Tensor, tile_start: torch.Tensor, piece: int | None):
    """Within-tile order by variable; rank, segment starts/tiles/variables."""
    device = index_sorted.device
    n = index_sorted.numel()
    ntiles = tile_start.numel() - 1
    tile_of = _tile_of(tile_start, n)
    pos = _arange(n, device)
    idx = index_sorted.to(torch.int64)
    perm = _stable_order([tile_of, idx])
    local = pos - tile_start[tile_of[perm]]  # position k in tile order (perm index k)
    rank = torch.empty(n, dtype=torch.int64, device=device)
    rank[perm] = local
    v = idx[perm]
    t = tile_of[perm]
    boundary = torch.ones(n, dtype=torch.bool, device=device)
    boundary[1:] = (v[1:] != v[:-1]) | (t[1:] != t[:-1])
    if piece is not None and n:
        run_id = torch.cumsum(boundary.to(torch.int64), 0) - 1
        run_begin = _flatnonzero(boundary)
        in_run = pos - run_begin[run_id]
        boundary |= (in_run % piece) == 0
    heads = _flatnonzero(boundary)
    seg_ptr = _counts_ptr(t[heads], ntiles, device)
    identity = bool(torch.equal(rank, pos - tile_start[tile_of]))
    return rank, identity, local[heads], seg_ptr, v[heads]


WAVE = 64


def _aligned_tiles(n: int, cap: int, device) -> torch.Tensor:
    """Tiles of exactly ``cap`` sorted positions (the last one may be short)."""
    if n == 0:
        return torch.zeros(1, dtype=torch.int64, device=device)
    starts = torch.arange(0, n, cap, dtype=torch.int64, device=device)
    return torch.cat((starts, torch.tensor([n], dtype=torch.int64, device=device)))


# Most rounds a tile may take in the occurrence-round reduction; groups with a
# more repetitive secondary slot (few variables) use the staged reduction.
MAX_ROUNDS = 8


def _occurrences(index_sorted: torch.Tensor, tile_start: torch.Tensor):
    """Occurrence index of each position's variable inside its tile (0 for the
    first position of the tile with that variable, in position order) and the
    number of rounds (1 + largest occurrence index) of each tile."""
    device = index_sorted.device
    n = index_sorted.numel()
    ntiles = tile_start.numel() - 1
    if n == 0:
        return torch.zeros(0, dtype=torch.int64, device=device), torch.zeros(ntiles, dtype=torch.int32, device=device)
    tile = _tile_of(tile_start, n)
    idx = index_sorted.to(torch.int64)
    order = _stable_order([tile, idx])
    kt, kv = tile[order], idx[order]
    first = torch.ones(n, dtype=torch.bool, device=device)
    first[1:] = (kt[1:] != kt[:-1]) | (kv[1:] != kv[:-1])
    pos = _arange(n, device)
    start = _flatnonzero(first)[torch.cumsum(first.to(torch.int64), 0) - 1]  # first position of each run
    occ = torch.empty(n, dtype=torch.int64, device=device)
    occ[order] = pos - start
    rounds = torch.zeros(ntiles, dtype=torch.int64, device=device)
    rounds.scatter_reduce_(0, tile, occ + 1, reduce="amax")
    return occ, rounds.to(torch.int32)


def _overflow(index_sorted: torch.Tensor, tile_start: torch.Tensor):
    """Overflow encoding of a dense secondary slot (wave kernel with staged outputs):
    rank 0 for the first position of each variable in its tile, else 1 + the slot of
    the position among the tile's later occurrences in (variable, position) order;
    and the number of later occurrences of each tile."""
    device = index_sorted.device
    n = index_sorted.numel()
    ntiles = tile_start.numel() - 1
    if n == 0:
        return torch.zeros(0, dtype=torch.int64, device=device), torch.zeros(ntiles, dtype=torch.int32, device=device)
    tile = _tile_of(tile_start, n)
    order = _stable_order([tile, index_sorted.to(torch.int64)])
    kt, kv = tile[order], index_sorted.to(torch.int64)[order]
    first = torch.ones(n, dtype=torch.bool, device=device)
    first[1:] = (kt[1:] != kt[:-1]) | (kv[1:] != kv[:-1])
    later = (~first).to(torch.int64)
    counts = torch.bincount(kt[~first], minlength=ntiles)
    before = torch.cumsum(counts, 0) - counts  # later occurrences in earlier tiles
    slot = torch.cumsum(later, 0) - 1 - before[kt]
    rank = torch.empty(n, dtype=torch.int64, device=device)
    rank[order] = torch.where(first, torch.zeros_like(slot), slot + 1)
    return rank, counts.to(torch.int32)


def _fast_lds(count_s: int, ds: int, dp: int, esize: int, per: int, out_entries: int, stage_entries: int) -> int:
    """LDS bytes of the staged-output wave kernel without the secondary x slice
    (csrc/normal.hip wave_lds): accumulator, overflow counts of the workgroup's
    tiles, one output region per wavefront and the overflow stage."""
    a16 = lambda b: -(-b // 16) * 16  # noqa: E731
    return (a16(count_s * ds * esize) + a16(per * 4) + (WAVE_THREADS // WAVE) * a16(out_entries * (dp * esize + 4))
            + a16(stage_entries * (ds * esize + 4)))


def _wave_pieces(index_sorted: torch.Tensor, tile_start: torch.Tensor, span: int):
    """Primary-slot pieces of the wave kernel: runs of one variable, also cut at
    every ``span``-position boundary of each tile (one wavefront covers 64
    threads x ITEMS consecutive positions)."""
    device = index_sorted.device
    n = index_sorted.numel()
    ntiles = tile_start.numel() - 1
    tile_of = _tile_of(tile_start, n)
    pos = _arange(n, device)
    local = pos - tile_start[tile_of]
    head = torch.ones(n, dtype=torch.bool, device=device)
    head[1:] = index_sorted[1:] != index_sorted[:-1]
    head |= (local % span) == 0
    heads = _flatnonzero(head)
    ends = torch.cat((heads[1:], torch.tensor([n], dtype=torch.int64, device=device))) - 1
    seg_ptr = _counts_ptr(tile_of[heads], ntiles, device)
    return local[heads], seg_ptr, index_sorted[heads].to(torch.int64), ends


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
    overflow: bool = True,
    force_primary: dict | None = None,
    split_primary: bool = True,
    wave_shapes: frozenset = frozenset(),
) -> StructurePlan:
    """Plan the deterministic reduction.

    ``layout``: [(count, dim, offset)] per variable type.
    ``groups``: [(rows, [type per slot], [int64 index tensor per slot])], every index
    tensor on the device the plan is built on (all groups on one device).
    ``overflow``: a wave group's dense secondary slot takes the overflow encoding
    (staged outputs, two barriers per tile) when it fits in LDS; otherwise, or
    when False, occurrence rounds (at most MAX_ROUNDS) or the staged reduction.
    """
    cap = tile_cap or THREADS * items_for(esize)
    types = [TypePlan(int(c), int(d), int(o)) for c, d, o in layout]
    device = next(torch.as_tensor(ix).device for _, _, idx in groups for ix in idx)
    gplans = []
    # staged wave outputs name positional partials with int32 ids
    ids_fit = sum(int(torch.as_tensor(idx[0]).numel()) for _, _, idx in groups) + 2 * WAVE_TILE < 2**31
    for gi, (rows, slot_types, slot_idx) in enumerate(groups):
        ns = len(slot_types)
        if not 1 <= ns <= 3:
            raise ValueError("the hip backend supports 1 to 3 variable slots per residual")
        slot_idx = [torch.as_tensor(ix).to(device=device, dtype=torch.int64).reshape(-1) for ix in slot_idx]
        count = int(slot_idx[0].numel())
        sizes = [types[t].count for t in slot_types]
        primary = int(np.argmax(sizes))
        if force_primary and gi in force_primary:
            primary = int(force_primary[gi])
        # sorted by the primary slot, then the other slots in slot order, then position
        keys = [slot_idx[primary]] + [slot_idx[s] for s in range(ns) if s != primary]
        order = _stable_order(keys) if count else torch.zeros(0, dtype=torch.int64, device=device)
        sorted_idx = [slot_idx[s][order] for s in range(ns)]
        dmax = max(types[t].dim for t in slot_types)
        shape = (int(rows), *(types[t].dim for t in slot_types))
        # (the wave kernel addresses each slot's x slice with 32-bit byte offsets)
        wave = (ns == 2 and shape in wave_shapes and count > 0
                and all(types[t].count * types[t].dim * esize < 2**32 for t in slot_types))
        # a group whose full tiles would leave most compute units idle is cut into
        # tiles of one observation per thread (the one-item tile kernels)
        small = tile_cap is None and not wave and -(-count // cap) < cu_count
        gcap = WAVE_THREADS * wave_items(esize) if wave else (THREADS if small else cap)
        tile_start = _aligned_tiles(count, gcap, device) if wave else _cut_tiles(sorted_idx[primary], gcap)
        ntiles = tile_start.numel() - 1
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
                    sorted_idx[s].to(torch.int32),
                    None if identity else rank.to(torch.int32),
                    seg_ptr.to(torch.int32),
                    seg_start.to(torch.int32),
                    seg_var.to(torch.int64),
                )
            )
            if wave and s == primary:
                slots[-1].piece_end = ends
        has_dense = any(sl.mode == REDUCE_DENSE for sl in slots)
        if has_dense and ntiles:
            # a wave-kernel workgroup (WAVE_THREADS = 512 threads of up to 256 VGPRs, two
            # wavefronts per SIMD) fills its compute unit's register file: one per CU
            nwg = min(ntiles, cu_count * (1 if wave else wgs_per_cu))
            per = -(-ntiles // nwg)
            nwg = -(-ntiles // per)
        else:
            nwg, per = ntiles, 1
        out_entries = stage_entries = 0
        sec = 1 - primary
        if wave and ntiles and slots[sec].mode == REDUCE_DENSE:
            occ, rounds = _occurrences(sorted_idx[sec], tile_start)
            use_rounds = bool(rounds.numel()) and int(rounds.max()) <= MAX_ROUNDS
            if use_rounds:
                slots[sec].rank = occ.to(torch.int32)
                slots[sec].rounds = rounds
            fits = False
            if overflow and ids_fit:
                # staged outputs: the most piece ends of one wavefront over a workgroup's tiles
                ps = slots[primary]
                span = WAVE * wave_items(esize)
                pieces = (ps.seg_ptr[1:] - ps.seg_ptr[:-1]).to(torch.int64)
                tile_of_piece = torch.repeat_interleave(_arange(ntiles, device), pieces)
                key = (tile_of_piece // per) * (WAVE_THREADS // WAVE) + ps.seg_start.to(torch.int64) // span
                out_entries = int(torch.bincount(key).max()) if key.numel() else 0
                ts, tp = types[slot_types[sec]], types[slot_types[primary]]
                fits = out_entries > 0 and _fast_lds(ts.count, ts.dim, tp.dim, esize, per, out_entries, 0) <= lds_limit
                if fits and not use_rounds:
                    # many repeats per tile: overflow encoding instead of the LDS stage
                    ov_rank, ov_counts = _overflow(sorted_idx[sec], tile_start)
                    stage_entries = max(1, int(ov_counts.max()))
                    fits = _fast_lds(ts.count, ts.dim, tp.dim, esize, per, out_entries, stage_entries) <= lds_limit
                    if fits:
                        slots[sec].rank = ov_rank.to(torch.int32)
                        slots[sec].rounds = ov_counts
            if not fits:
                out_entries = stage_entries = 0
        gplans.append(
            GroupPlan(count, int(rows), primary, order, tile_start.to(torch.int32), gcap, nwg, per, slots,
                      kernel=1 if wave else 0, out_entries=out_entries, stage_entries=stage_entries)
        )
    # dense blocks first (ids b*count + v), then sparse partials
    for g in gplans:
        for sl in g.slots:
            if sl.mode == REDUCE_DENSE:
                sl.dense_block = types[sl.type].dense_blocks
                types[sl.type].dense_blocks += g.nwg
                sl.seg_target = sl.seg_var.to(torch.int32)
    pairs = [[] for _ in types]
    next_id = [t.dense_blocks * t.count for t in types]
    for g in gplans:
        for sl in g.slots:
            if sl.mode == REDUCE_SPARSE and g.kernel == 0:
                nseg = sl.seg_var.numel()
                ids = next_id[sl.type] + _arange(nseg, device)
                next_id[sl.type] += nseg
                sl.seg_target = ids.to(torch.int32)
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
                tile_of = _tile_of(sl.seg_ptr.to(torch.int64), sl.seg_start.numel())
                ids = sl.pos_base + g.tile_start[tile_of].to(torch.int64) + sl.seg_start.to(torch.int64)
            sl.seg_target = ids.to(torch.int64)
            pairs[sl.type].append((sl.seg_var, ids))
    # a variable whose only contribution is one primary wave piece is written directly
    for t, ty in enumerate(types):
        contrib = torch.zeros(ty.count, dtype=torch.int64, device=device)
        if ty.dense_blocks:
            contrib += 2
        for var, _ in pairs[t]:
            contrib += torch.bincount(var, minlength=ty.count)
        direct = torch.zeros(ty.count, dtype=torch.bool, device=device)
        for g in gplans:
            if g.kernel == 1:
                sl = g.slots[g.primary]
                if sl.type == t:
                    cand = torch.zeros(ty.count, dtype=torch.bool, device=device)
                    cand[sl.seg_var] = True
                    direct |= cand & (contrib == 1)
        if bool(direct.any()):
            ty.direct = direct.to(torch.uint8)
            ty.finish_list = _flatnonzero(~direct).to(torch.int32)
            keep = []
            for var, ids in pairs[t]:
                m = ~direct[var]
                keep.append((var[m], ids[m]))
            pairs[t] = keep
            for g in gplans:
                if g.kernel == 1 and g.slots[g.primary].type == t:
                    sl = g.slots[g.primary]
                    sl.seg_target = torch.where(direct[sl.seg_var], torch.full_like(sl.seg_target, -1), sl.seg_target)
    for g in gplans:
        for sl in g.slots:
            if sl.seg_target is not None and sl.seg_target.dtype != torch.int32:
                sl.seg_target = sl.seg_target.to(torch.int32)
    for t, ty in enumerate(types):
        ty.sparse_partials = next_id[t] - ty.dense_blocks * ty.count
        if next_id[t] > 2**31 - 1:
            raise ValueError("partial ids exceed the 32-bit range")
        if pairs[t]:
            var = torch.cat([p[0] for p in pairs[t]])
            ids = torch.cat([p[1] for p in pairs[t]])
            o = _stable_order([var, ids])
            ty.csr_idx = ids[o].to(torch.int32)
            ty.csr_ptr = _counts_ptr(var, ty.count, device).to(torch.int32)
        else:
            ty.csr_idx = torch.zeros(0, dtype=torch.int32, device=device)
            ty.csr_ptr = torch.zeros(ty.count + 1, dtype=torch.int32, device=device)
    return StructurePlan(types, gplans, esize, cap)


# Variable-major accumulation (csrc/major.hip) pays off for a type with few
# variables observed many times each: its blocks then need no per-tile partials.
MAJOR_MAX_VARIABLES = 65536
MAJOR_MIN_MEAN = 64


def major_eligible(variables: int, observations: int) -> bool:
    return 0 < variables <= MAJOR_MAX_VARIABLES and observations >= MAJOR_MIN_MEAN * variables \
        and observations < 2**31


def major_order(index: torch.Tensor, count: int):
    """Observations sorted by variable (stable) and each variable's range: (order, ptr), int32."""
    index = torch.as_tensor(index).to(torch.int64).reshape(-1)
    order = torch.sort(index, stable=True).indices
    return order.to(torch.int32), _counts_ptr(index, count, index.device).to(torch.int32)


def major_edges(index_c: torch.Tensor, index_p: torch.Tensor, count_c: int):
    """Observations sorted by (camera, point, position) and their (camera, point) runs:
    (order, edge_ptr [E+1], var_edge_ptr [count_c+1], edge_point [E]), all int32."""
    ic = torch.as_tensor(index_c).to(torch.int64).reshape(-1)
    ip = torch.as_tensor(index_p).to(torch.int64).reshape(-1).to(ic.device)
    n = ic.numel()
    order = _stable_order([ic, ip])
    sc, sp = ic[order], ip[order]
    head = torch.ones(n, dtype=torch.bool, device=ic.device)
    if n:
        head[1:] = (sc[1:] != sc[:-1]) | (sp[1:] != sp[:-1])
    starts = _flatnonzero(head)
    edge_ptr = torch.cat((starts, torch.tensor([n], dtype=torch.int64, device=ic.device)))
    var_edge_ptr = _counts_ptr(sc[starts], count_c, ic.device)
    return (order.to(torch.int32), edge_ptr.to(torch.int32), var_edge_ptr.to(torch.int32),
            sp[starts].to(torch.int32))
