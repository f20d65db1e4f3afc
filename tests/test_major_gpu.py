"""Variable-major block accumulations (csrc/major.hip) against numpy and the tile kernels."""

import ctypes

import numpy as np
import pytest
import torch

from tests.helpers import rel_err
from tests.test_schur_gpu import build
from torchls_amd import _lib, generators
from torchls_amd.hip_diagonal import HipDiagonalPlan
from torchls_amd.plan import major_order
from torchls_amd.schur import HipSchurPlan

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rows,dim", [(2, 6), (2, 3), (6, 6)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_major_gram_matches_numpy(device, rows, dim, dtype):
    print(f"major gram rows={rows} dim={dim} dtype={dtype}", flush=True)
    rng = np.random.default_rng(rows * 10 + dim)
    count, n = 9, 3000
    idx = rng.integers(0, count - 1, n)  # the last variable has no observations
    blocks = rng.standard_normal((n, rows, dim)).astype(np.float64 if dtype == torch.float64 else np.float32)
    order, ptr = major_order(idx, count)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)  # noqa: E731
    b, o, p = dev(blocks), dev(order), dev(ptr)
    out = torch.full((count, dim, dim), np.nan, dtype=torch.float64, device=device)
    lib = _lib.load()
    _lib.check(lib.tlh_major_gram(b.data_ptr(), _lib.dtype_code(b.dtype), rows, dim, o.data_ptr(), p.data_ptr(),
                                  count, out.data_ptr(), _lib.stream_handle(device)), "tlh_major_gram")
    bd = blocks.astype(np.float64)
    want = np.zeros((count, dim, dim))
    np.add.at(want, idx, np.einsum("nri,nrj->nij", bd, bd))
    got = out.cpu().numpy()
    assert np.all(got[count - 1] == 0.0)
    assert rel_err(got, want) <= 1e-12
    # fixed summation order: a second launch is bitwise identical
    again = torch.empty_like(out)
    _lib.check(lib.tlh_major_gram(b.data_ptr(), _lib.dtype_code(b.dtype), rows, dim, o.data_ptr(), p.data_ptr(),
                                  count, again.data_ptr(), _lib.stream_handle(device)), "tlh_major_gram")
    assert torch.equal(out, again)


def test_major_gram_rejects_unknown_shape(device):
    lib = _lib.load()
    dummy = torch.zeros(8, dtype=torch.float64, device=device)
    idx = torch.zeros(8, dtype=torch.int32, device=device)
    code = lib.tlh_major_gram(dummy.data_ptr(), _lib.TLH_F64, 2, 5, idx.data_ptr(), idx.data_ptr(), 1,
                              dummy.data_ptr(), _lib.stream_handle(device))
    assert code == 1002  # TLH_ENOTFOUND: no kernel for 2x5 blocks


@pytest.mark.parametrize("priors", [False, True])
def test_diagonal_major_matches_tile_path(device, priors):
    print(f"diagonal major priors={priors}", flush=True)
    P = build(generators.bundle_adjustment(16, 500, 8, seed=23), device, priors)
    J = P._analyze().linearize(P.parameters).jacobian
    plan = HipDiagonalPlan(J)
    # cameras qualify only when one residual group observes them (no prior)
    assert (0 in plan.major) == (not priors)
    packed = plan._pack(J)
    tile = plan.diagonal_blocks_f64(packed)
    major = plan.diagonal_blocks_f64(packed, J)
    for a, b in zip(major, tile):
        assert rel_err(a.cpu().numpy(), b.cpu().numpy()) <= 1e-12
        assert torch.equal(a, a.transpose(1, 2))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_schur_precond_major_matches_tile_path(device, dtype):
    print(f"schur precond major dtype={dtype}", flush=True)
    P = build(generators.bundle_adjustment(16, 500, 8, seed=24), device)
    an = P._analyze()
    lin = an.linearize(P.parameters)
    diag = tuple(d.double() for d in lin.jacobian.diagonal_blocks())
    plan = HipSchurPlan(an, "points", device, dtype=torch.float64)
    system = plan.linearize(lin.jacobian, diag).damped(3e-2)
    assert plan.major is not None
    major = [b.clone() for b in system.preconditioner_blocks()]
    saved, plan.major = plan.major, None
    try:
        tile = system.preconditioner_blocks()
    finally:
        plan.major = saved
    for a, b in zip(major, tile):
        assert rel_err(a.cpu().numpy(), b.cpu().numpy()) <= 1e-12
    assert torch.equal(major[0], system.preconditioner_blocks()[0])


def test_schur_precond_major_repeated_edges(device):
    """A (camera, point) pair observed twice is one edge: W_e sums both observations."""
    print("schur precond major repeated edges", flush=True)
    prob = generators.bundle_adjustment(8, 300, 6, seed=25)
    # duplicate every 7th observation so those edges hold two observations
    dup = np.arange(0, len(prob.cam_idx), 7)
    for name in ("cam_idx", "pt_idx", "observations"):
        arr = getattr(prob, name)
        setattr(prob, name, np.concatenate([arr, arr[dup]]))
    P = build(prob, device)
    an = P._analyze()
    lin = an.linearize(P.parameters)
    diag = tuple(d.double() for d in lin.jacobian.diagonal_blocks())
    plan = HipSchurPlan(an, "points", device, dtype=torch.float64)
    system = plan.linearize(lin.jacobian, diag).damped(1e-2)
    assert plan.major is not None
    order, edge_ptr, _, _ = (t.cpu().numpy() for t in plan.major)
    assert np.any(np.diff(edge_ptr) == 2)
    major = system.preconditioner_blocks()[0].clone()
    saved, plan.major = plan.major, None
    try:
        tile = system.preconditioner_blocks()[0]
    finally:
        plan.major = saved
    assert rel_err(major.cpu().numpy(), tile.cpu().numpy()) <= 1e-12


@pytest.mark.parametrize("rows,dim", [(2, 6), (2, 3), (6, 6)])
def test_major_rmatvec_matches_numpy(device, rows, dim):
    print(f"major rmatvec rows={rows} dim={dim}", flush=True)
    rng = np.random.default_rng(40 + rows + dim)
    count, n = 7, 2500
    idx = rng.integers(0, count, n)
    blocks = rng.standard_normal((n, rows, dim))
    v = rng.standard_normal(n * rows)
    order, ptr = major_order(idx, count)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)  # noqa: E731
    b, vv, o, p = dev(blocks), dev(v), dev(order), dev(ptr)
    out = torch.full((count, dim), np.nan, dtype=torch.float64, device=device)
    lib = _lib.load()
    _lib.check(lib.tlh_major_rmatvec(b.data_ptr(), _lib.TLH_F64, vv.data_ptr(), _lib.TLH_F64, rows, dim, o.data_ptr(),
                                     p.data_ptr(), count, out.data_ptr(), _lib.stream_handle(device)),
               "tlh_major_rmatvec")
    want = np.zeros((count, dim))
    np.add.at(want, idx, np.einsum("nrc,nr->nc", blocks, v.reshape(n, rows)))
    assert rel_err(out.cpu().numpy(), want) <= 1e-12


def test_gradient_major_matches_tile_path(device):
    print("gradient major vs tile", flush=True)
    from torchls_amd.hip_ops import LinearizationKernels

    P = build(generators.bundle_adjustment(16, 500, 8, seed=26), device)
    lin = P._analyze().linearize(P.parameters)
    kernels = LinearizationKernels(lin.jacobian)
    assert 0 in kernels.diag_plan.major
    r = lin.residual.to(torch.float64).contiguous()
    g_major = kernels.gradient(lin.jacobian, r)
    saved, kernels.diag_plan.major = kernels.diag_plan.major, {}
    try:
        g_tile = kernels.gradient(lin.jacobian, r)
    finally:
        kernels.diag_plan.major = saved
    assert rel_err(g_major.cpu().numpy(), g_tile.cpu().numpy()) <= 1e-12
    J = lin.jacobian
    dense = torch.zeros(J.shape, dtype=torch.float64)
    row = 0
    for g in J.groups:
        n, rows = g.blocks[0].shape[:2]
        for b, t, vi in zip(g.blocks, g.type_indices, g.variable_indices):
            ent = J.layout[t]
            bd = b.double().cpu()
            vic = vi.reshape(-1).cpu().numpy()
            for k in range(n):
                c0 = ent.offset + int(vic[k]) * ent.dim
                dense[row + k * rows: row + (k + 1) * rows, c0: c0 + ent.dim] += bd[k]
        row += n * rows
    want = (dense.T @ r.cpu()).numpy()
    assert rel_err(g_major.cpu().numpy(), want) <= 1e-12


@pytest.mark.parametrize("src_dt,dst_dt", [(torch.float64, torch.float32), (torch.float32, torch.float64),
                                           (torch.float64, torch.float64)])
def test_copy_cast_gathers_and_scatters(device, src_dt, dst_dt):
    print(f"copy_cast {src_dt} -> {dst_dt}", flush=True)
    lib = _lib.load()
    g = torch.Generator().manual_seed(7)
    src = torch.randn(1000, generator=g, dtype=torch.float64).to(src_dt).to(device)
    idx = torch.randperm(1000, generator=g)[:300].to(device)
    out = torch.full((300,), np.nan, dtype=dst_dt, device=device)
    stream = _lib.stream_handle(device)
    # gather with a scale
    _lib.check(lib.tlh_copy_cast(src.data_ptr(), _lib.dtype_code(src_dt), idx.data_ptr(), 0, out.data_ptr(),
                                 _lib.dtype_code(dst_dt), None, 0, 300, -1.0, stream), "tlh_copy_cast")
    want = (-src.double()[idx]).to(dst_dt)
    assert torch.equal(out, want)
    # scatter into an offset-free destination, then a contiguous copy at an offset
    dst = torch.full((1000,), np.nan, dtype=dst_dt, device=device)
    _lib.check(lib.tlh_copy_cast(out.data_ptr(), _lib.dtype_code(dst_dt), None, 0, dst.data_ptr(),
                                 _lib.dtype_code(dst_dt), idx.data_ptr(), 0, 300, 1.0, stream), "tlh_copy_cast")
    assert torch.equal(dst[idx], out)
    _lib.check(lib.tlh_copy_cast(src.data_ptr(), _lib.dtype_code(src_dt), None, 10, dst.data_ptr(),
                                 _lib.dtype_code(dst_dt), None, 500, 100, 1.0, stream), "tlh_copy_cast")
    assert torch.equal(dst[500:600], src[10:110].double().to(dst_dt))
