"""CPU checks of the tile plan: executing its data flow reproduces J^T J x."""

import numpy as np
import pytest

import torch

from oracle.sparse import random_jacobian
from tests.plan_emulation import emulate_normal_matvec, pack_numpy
from torchls_amd.plan import REDUCE_DENSE, build_structure


def _t(a):
    return torch.as_tensor(np.asarray(a, dtype=np.int64))

SHAPES = [
    ([(7, 6), (300, 3), (5, 1)], [(1500, 2, [0, 1]), (40, 3, [1]), (200, 6, [0, 0]), (50, 2, [2, 1, 2])]),
    ([(3, 7)], [(100, 4, [0, 0]), (1, 1, [0])]),
    ([(50, 6), (20, 3)], [(900, 2, [1, 0])]),
]


@pytest.mark.parametrize("shape", range(len(SHAPES)))
@pytest.mark.parametrize("cap,piece,lds", [(64, 4, 65536), (1024, 32, 65536), (37, 5, 65536), (256, 3, 100)])
def test_plan_emulation_matches_oracle(shape, cap, piece, lds):
    rng = np.random.default_rng(shape * 100 + cap)
    layout, groups = SHAPES[shape]
    J = random_jacobian(rng, layout, groups)
    plan = build_structure(
        J.layout, [(g.rows, g.types, [_t(i) for i in g.indices]) for g in J.groups],
        esize=8, tile_cap=cap, piece=piece, lds_limit=lds,
    )
    for g in plan.groups:
        assert int(g.tile_start[-1]) == g.count
        assert bool((torch.diff(g.tile_start) <= cap).all())
        for s in g.slots:
            if s.mode != REDUCE_DENSE:
                assert s.seg_start.numel() == s.seg_target.numel()
    x = rng.standard_normal(J.n)
    y = emulate_normal_matvec(plan, pack_numpy(plan, [g.blocks for g in J.groups]), x, 0.25)
    dense = J.to_dense()
    ref = dense.T @ (dense @ x) + 0.25 * x
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()


def test_primary_is_largest_type_and_tiles_align():
    rng = np.random.default_rng(3)
    J = random_jacobian(rng, [(10, 6), (400, 3)], [(3000, 2, [0, 1])])
    plan = build_structure(J.layout, [(g.rows, g.types, [_t(i) for i in g.indices]) for g in J.groups], esize=4,
                           tile_cap=256)
    g = plan.groups[0]
    assert g.primary == 1 and g.slots[1].rank is None and g.slots[0].rank is not None
    prim = g.slots[1].index.numpy()
    # every tile boundary falls between two different points (runs fit in a tile)
    for b in g.tile_start[1:-1].tolist():
        assert prim[b] != prim[b - 1]
    assert g.slots[0].mode == REDUCE_DENSE


def test_major_order_and_edges_match_brute_force():
    from torchls_amd.plan import major_edges, major_eligible, major_order

    rng = np.random.default_rng(5)
    n, nc, npt = 700, 6, 40
    ic = rng.integers(0, nc - 1, n)  # camera nc - 1 is never observed
    ip = rng.integers(0, npt, n)
    order, ptr = major_order(_t(ic), nc)
    assert order.dtype == torch.int32 and ptr.dtype == torch.int32
    order, ptr = order.numpy(), ptr.numpy()
    for v in range(nc):
        members = order[ptr[v]:ptr[v + 1]]
        assert np.array_equal(members, np.flatnonzero(ic == v))  # stable: increasing positions
    order, edge_ptr, var_edge_ptr, edge_point = (a.numpy() for a in major_edges(_t(ic), _t(ip), nc))
    assert sorted(order.tolist()) == list(range(n))
    assert var_edge_ptr[-1] == edge_point.size == edge_ptr.size - 1 and edge_ptr[-1] == n
    assert var_edge_ptr[nc - 1] == var_edge_ptr[nc]
    for v in range(nc):
        pairs = sorted({int(p) for p in ip[ic == v]})
        edges = range(var_edge_ptr[v], var_edge_ptr[v + 1])
        assert [int(edge_point[e]) for e in edges] == pairs
        for e in edges:
            obs = order[edge_ptr[e]:edge_ptr[e + 1]]
            assert np.all(ic[obs] == v) and np.all(ip[obs] == edge_point[e])
            assert np.array_equal(obs, np.flatnonzero((ic == v) & (ip == edge_point[e])))
    assert major_eligible(200, 500_000) and not major_eligible(50_000, 500_000) and not major_eligible(0, 10)


def test_operator_descriptor_cache():
    """DeviceStructure.operator returns one descriptor per (dtype, block and workspace
    addresses): the same object for the same addresses, a new one carrying the new
    addresses otherwise (the descriptor holds addresses only; no GPU call is made)."""
    from torchls_amd.hip_normal import DeviceStructure

    rng = np.random.default_rng(5)
    layout, groups = SHAPES[2]
    J = random_jacobian(rng, layout, groups)
    plan = build_structure(J.layout, [(g.rows, g.types, [_t(i) for i in g.indices]) for g in J.groups], esize=4)
    dev = DeviceStructure(plan, torch.device("cpu"))
    a, b = torch.zeros(64), torch.zeros(64)
    parts = [torch.zeros(8) for _ in plan.types]
    op1 = dev.operator(torch.float32, [a], parts)
    assert dev.operator(torch.float32, [a], parts) is op1
    op2 = dev.operator(torch.float32, [b], parts)
    assert op2 is not op1
    assert op1.groups[0].packed == a.data_ptr() and op2.groups[0].packed == b.data_ptr()
    assert op2.types[0].partials == parts[0].data_ptr()
    op3 = dev.operator(torch.float64, [a], parts)
    assert op3 is not op1 and op3.dtype != op1.dtype
    assert op1.ngroups == len(plan.groups) and op1.size == plan.size
