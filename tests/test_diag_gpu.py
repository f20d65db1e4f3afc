"""GPU parity of diagonal blocks, block inverses and the float64 LM reductions."""

import numpy as np
import pytest
import torch

from oracle.ba import huber_reduction, huber_value
from oracle.sparse import random_jacobian
from tests.helpers import ba_jacobian, rel_err, to_block_jacobian
from torchls_amd import _lib, generators
from torchls_amd.hip_diagonal import HipBlockJacobi, HipDiagonalPlan, LinAlgError
from torchls_amd.hip_ops import LinearizationKernels, Reduction, trial_objective

pytestmark = pytest.mark.gpu

SHAPES = [
    ([(7, 6), (300, 3), (5, 1)], [(1500, 2, [0, 1]), (40, 3, [1]), (200, 6, [0, 0]), (50, 2, [2, 1, 2])]),
    ([(3, 7)], [(100, 4, [0, 0]), (1, 1, [0])]),
    ([(40, 6)], [(3000, 6, [0, 0])]),
]


@pytest.mark.parametrize("shape", range(len(SHAPES)))
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 1e-5)])
def test_diagonal_blocks(device, shape, dtype, tol):
    print(f"diag shape={shape} dtype={dtype}", flush=True)
    rng = np.random.default_rng(shape)
    J = random_jacobian(rng, *SHAPES[shape])
    # force aliases on duplicated-type slots for part of the observations
    for g in J.groups:
        for s in range(1, len(g.types)):
            if g.types[s] == g.types[0]:
                g.indices[s][::3] = g.indices[0][::3]
    if dtype == torch.float32:
        for g in J.groups:
            g.blocks = [b.astype(np.float32).astype(np.float64) for b in g.blocks]
    ref = J.diagonal_blocks()
    BJ = to_block_jacobian(J, device, dtype)
    out = HipDiagonalPlan(BJ).diagonal_blocks(BJ)
    for o, r in zip(out, ref):
        assert o.dtype == dtype
        assert rel_err(o.cpu().numpy(), r) <= tol
        assert torch.equal(o, o.transpose(-1, -2))


def test_block_inverse(device):
    rng = np.random.default_rng(0)
    A = rng.standard_normal((500, 6, 6))
    D = A @ np.swapaxes(A, 1, 2)
    B = rng.standard_normal((300, 3, 3))
    E = B @ np.swapaxes(B, 1, 2)
    from torchls_amd.sparse import VariableLayout
    from tests.helpers import _Dim

    layout = (VariableLayout(_Dim(6), torch.arange(500), 0), VariableLayout(_Dim(3), torch.arange(300), 3000))
    pre = HipBlockJacobi(layout, (torch.tensor(D, device=device), torch.tensor(E, device=device)), 0.5)
    r = rng.standard_normal(3900)
    z = pre(torch.tensor(r, device=device)).cpu().numpy()
    zr = np.concatenate((np.linalg.solve(D + 0.5 * np.eye(6), r[:3000].reshape(500, 6)[..., None])[..., 0].ravel(),
                         np.linalg.solve(E + 0.5 * np.eye(3), r[3000:].reshape(300, 3)[..., None])[..., 0].ravel()))
    assert rel_err(z, zr) <= 1e-10
    with pytest.raises(LinAlgError):
        HipBlockJacobi(layout, (torch.tensor(D, device=device), torch.tensor(-E, device=device)), 0.0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_gradient_sqnorm_matvec(device, dtype):
    prob = generators.bundle_adjustment(16, 500, 8, seed=11)
    J, r = ba_jacobian(prob, delta=1e-2)
    if dtype == torch.float32:
        for g in J.groups:
            g.blocks = [b.astype(np.float32).astype(np.float64) for b in g.blocks]
    BJ = to_block_jacobian(J, device, dtype)
    k = LinearizationKernels(BJ)
    rt = torch.tensor(r.reshape(-1), device=device, dtype=dtype)
    gr = k.gradient(BJ, rt)
    ref_r = rt.double().cpu().numpy()
    assert gr.dtype == torch.float64
    assert rel_err(gr.cpu().numpy(), J.rmatvec(ref_r)) <= 1e-12
    x = np.random.default_rng(1).standard_normal(J.n)
    sq = float(k.squared_matvec_norm(BJ, torch.tensor(x, device=device)))
    assert abs(sq - J.squared_matvec_norm(x)) <= 1e-12 * J.squared_matvec_norm(x)
    y = k.matvec(BJ, torch.tensor(x, device=device, dtype=dtype))
    assert rel_err(y.cpu().numpy(), J.matvec(x)) <= (1e-12 if dtype == torch.float64 else 1e-6)


def test_trial_objective(device):
    rng = np.random.default_rng(4)
    before = rng.standard_normal((2000, 2)) * 0.02
    after = before + rng.standard_normal((2000, 2)) * 0.005
    red = Reduction(device)
    w = rng.random(2000)
    for delta, weights in [(None, None), (1e-2, None), (1e-2, w), (None, w)]:
        trial_objective(red, 0, torch.tensor(before, device=device), torch.tensor(after, device=device), 2000, 2,
                        huber_delta=delta, weights=None if weights is None else torch.tensor(weights, device=device))
        so, sn = (before**2).sum(-1), (after**2).sum(-1)
        diff = ((before - after) * (before + after)).sum(-1)
        val = sn if delta is None else huber_value(sn, delta)
        chg = diff if delta is None else huber_reduction(so, sn, diff, delta)
        if weights is not None:
            val, chg = val * weights, chg * weights
        got = red.out[0].cpu().numpy()
        assert abs(got[0] - val.sum()) <= 1e-12 * abs(val.sum())
        assert abs(got[1] - chg.sum()) <= 1e-12 * abs(chg).sum()


@pytest.mark.parametrize("n", [1, 1000, 300_001])
def test_vector_stats(device, n):
    """tlh_vector_stats = (a.b, b.b, max|b|, max|a|) against float64 numpy, float64 a and
    float32 b; a NaN entry makes a maximum NaN, an infinite one infinite; bitwise repeatable."""
    print(f"vector stats n={n}", flush=True)
    from torchls_amd.hip_ops import Reduction

    rng = np.random.default_rng(n)
    a = rng.standard_normal(n)
    b = rng.standard_normal(n).astype(np.float32)
    red = Reduction(device)
    lib = _lib.load()

    def stats(x, y):
        out = torch.empty(4, dtype=torch.float64, device=device)
        _lib.check(lib.tlh_vector_stats(None if x is None else x.data_ptr(), 0 if x is None else _lib.dtype_code(x.dtype),
                                        None if y is None else y.data_ptr(), 0 if y is None else _lib.dtype_code(y.dtype),
                                        n, out.data_ptr(), red.partials.data_ptr(), red.counter.data_ptr(),
                                        _lib.stream_handle(device)), "tlh_vector_stats")
        return out.cpu().numpy()

    ta, tb = torch.tensor(a, device=device), torch.tensor(b, device=device)
    got = stats(ta, tb)
    b64 = b.astype(np.float64)
    assert abs(got[0] - a @ b64) <= 1e-12 * (np.abs(a) @ np.abs(b64))
    assert abs(got[1] - b64 @ b64) <= 1e-12 * (b64 @ b64)
    assert got[2] == np.abs(b64).max() and got[3] == np.abs(a).max()
    assert np.array_equal(got, stats(ta, tb))
    only = stats(None, tb)
    assert only[1] == got[1] and only[2] == got[2] and only[3] == 0.0
    bad_b, bad_a = b.copy(), a.copy()
    bad_b[n // 2] = np.nan
    bad_a[n - 1] = np.inf
    got = stats(torch.tensor(bad_a, device=device), torch.tensor(bad_b, device=device))
    assert np.isnan(got[1]) and np.isnan(got[2]) and got[3] == np.inf
