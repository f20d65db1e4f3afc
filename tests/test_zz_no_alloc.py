"""The no-allocation Cholesky solve (library-owned workspace) through the C ABI
gives the same result as the workspace-argument entry point, on two calls in a
row with the second system smaller (the array keeps the first call's contents)."""

import numpy as np
import pytest
import torch

from tests.helpers import rel_err
from torchls_amd import _lib
from torchls_amd.cholesky import cholesky_f64

pytestmark = pytest.mark.gpu


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n + 3))
    return A @ A.T + n * 1e-3 * np.eye(n), rng.standard_normal(n)


def _solve_with_work(lib, L, b, stream):
    n = b.numel()
    x = b.clone()
    work = torch.empty(2 * n, dtype=torch.float64, device=b.device)
    status = torch.zeros(1, dtype=torch.int32, device=b.device)
    _lib.check(lib.tlh_cholesky_solve_blocks_f64(L.data_ptr(), n, n, x.data_ptr(), work.data_ptr(),
                                                 status.data_ptr(), stream), "tlh_cholesky_solve_blocks_f64")
    return x, status


def test_noalloc_solve_matches_workspace_solve(device):
    lib = _lib.load()
    stream = _lib.stream_handle(device)
    for n, seed in ((1000, 11), (300, 12)):  # second call smaller than the first
        print(f"no-alloc cholesky solve n={n}", flush=True)
        M, b = _spd(n, seed)
        L = cholesky_f64(torch.tensor(M, device=device))
        bt = torch.tensor(b, device=device)
        ref, ref_status = _solve_with_work(lib, L, bt, stream)
        x = bt.clone()
        status = torch.zeros(1, dtype=torch.int32, device=device)
        _lib.check(lib.tlh_cholesky_solve_blocks_f64_noalloc(L.data_ptr(), n, n, x.data_ptr(), status.data_ptr(),
                                                             stream), "tlh_cholesky_solve_blocks_f64_noalloc")
        torch.cuda.synchronize()
        assert int(status.item()) == 0 and int(ref_status.item()) == 0
        assert torch.equal(x, ref)  # same kernels, same order: bitwise equal
        assert rel_err(x.cpu().numpy(), np.linalg.solve(M, b)) <= 1e-10


def test_noalloc_solve_rejects_oversized_systems(device):
    lib = _lib.load()
    b = torch.zeros(8, dtype=torch.float64, device=device)
    status = torch.zeros(1, dtype=torch.int32, device=device)
    # argument check only: nothing is launched for n above the 8192 maximum
    rc = lib.tlh_cholesky_solve_blocks_f64_noalloc(b.data_ptr(), 8193, 8193, b.data_ptr(), status.data_ptr(),
                                                   _lib.stream_handle(device))
    assert rc == _lib.TLH_EINVAL
