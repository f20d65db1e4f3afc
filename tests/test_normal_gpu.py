"""GPU parity of the HIP normal matvec against the float64 oracle."""

import numpy as np
import pytest
import torch

from oracle.sparse import random_jacobian
from tests.helpers import ba_jacobian, rel_err, to_block_jacobian
from torchls_amd import generators
from torchls_amd.hip_normal import HipNormalPlan

pytestmark = pytest.mark.gpu

SHAPES = [
    ([(7, 6), (300, 3), (5, 1)], [(1500, 2, [0, 1]), (40, 3, [1]), (200, 6, [0, 0]), (50, 2, [2, 1, 2])]),
    ([(3, 7)], [(100, 4, [0, 0]), (1, 1, [0])]),
    ([(50, 6), (20, 3)], [(900, 2, [1, 0])]),
    ([(40, 6)], [(3000, 6, [0, 0]), (1, 6, [0])]),
]


@pytest.mark.parametrize("shape", range(len(SHAPES)))
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 1e-5)])
def test_random_structures(device, shape, dtype, tol):
    print(f"normal matvec random shape={shape} dtype={dtype}", flush=True)
    rng = np.random.default_rng(10 + shape)
    layout, groups = SHAPES[shape]
    J = random_jacobian(rng, layout, groups)
    if dtype == torch.float32:
        for g in J.groups:
            g.blocks = [b.astype(np.float32).astype(np.float64) for b in g.blocks]
    x = rng.standard_normal(J.n)
    ref = J.normal_matvec(x, 0.125)
    BJ = to_block_jacobian(J, device, dtype)
    op = HipNormalPlan(BJ).bind(BJ)
    y = op.normal_matvec(torch.tensor(x, dtype=dtype, device=device), 0.125)
    assert y.dtype == dtype and y.shape == (J.n,)
    assert rel_err(y.cpu().numpy(), ref) <= tol


# (2000, 6000, 4) has many cameras per tile, so the camera slot is reduced in
# occurrence rounds; with fewer cameras a camera repeats too often in a tile.
# overflow=True: piece ends are staged in LDS, and the repetitive camera slots
# take the overflow encoding (first occurrences added directly, later ones
# staged and summed per camera run).  overflow=False: the repetitive camera
# slots stage camera segments in LDS and piece ends are written directly.
@pytest.mark.parametrize("overflow", [True, False])
@pytest.mark.parametrize("n_cams,n_pts,k", [(16, 500, 8), (200, 50_000, 10), (2000, 6000, 4)])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 1e-5)])
def test_bundle_adjustment(device, n_cams, n_pts, k, dtype, tol, overflow):
    print(f"normal matvec BA cams={n_cams} pts={n_pts} dtype={dtype} overflow={overflow}", flush=True)
    prob = generators.bundle_adjustment(n_cams, n_pts, k, seed=7)
    J, _ = ba_jacobian(prob, delta=1e-2)
    if dtype == torch.float32:
        for g in J.groups:
            g.blocks = [b.astype(np.float32).astype(np.float64) for b in g.blocks]
    rng = np.random.default_rng(1)
    x = rng.standard_normal(J.n)
    ref = J.normal_matvec(x, 1e-3)
    BJ = to_block_jacobian(J, device, dtype)
    plan = HipNormalPlan(BJ, overflow=overflow)
    g = plan.plan.groups[0]
    assert g.kernel == 1
    if overflow:
        assert g.out_entries > 0 and (g.stage_entries > 0) == (n_cams != 2000)
    else:
        assert g.out_entries == 0 and (g.slots[0].rounds is not None) == (n_cams == 2000)
    op = plan.bind(BJ)
    xt = torch.tensor(x, dtype=dtype, device=device)
    y1 = op.normal_matvec(xt, 1e-3)
    y2 = op.normal_matvec(xt, 1e-3)
    assert torch.equal(y1, y2), "normal matvec must be bitwise deterministic"
    assert rel_err(y1.cpu().numpy(), ref) <= tol


def test_mixed_precision_storage(device):
    """float64 blocks packed to float32 (dtype=float32) match a float32 copy."""
    prob = generators.bundle_adjustment(16, 500, 8, seed=3)
    J, _ = ba_jacobian(prob)
    BJ64 = to_block_jacobian(J, device, torch.float64)
    BJ32 = to_block_jacobian(J, device, torch.float32)
    x = torch.randn(J.n, dtype=torch.float32, device=device)
    a = HipNormalPlan(BJ64, dtype=torch.float32).bind(BJ64).normal_matvec(x, 0.5)
    b = HipNormalPlan(BJ32).bind(BJ32).normal_matvec(x, 0.5)
    assert torch.equal(a, b)


def test_errors(device):
    rng = np.random.default_rng(0)
    J = random_jacobian(rng, [(4, 3)], [(10, 2, [0])])
    BJ = to_block_jacobian(J, device, torch.float64)
    op = HipNormalPlan(BJ).bind(BJ)
    with pytest.raises(ValueError):
        op.normal_matvec(torch.zeros(J.n + 1, dtype=torch.float64, device=device))
    with pytest.raises(ValueError):
        op.normal_matvec(torch.zeros(J.n, dtype=torch.float32, device=device))
    cpu = to_block_jacobian(J, "cpu", torch.float64)
    with pytest.raises(ValueError):
        HipNormalPlan(cpu)
