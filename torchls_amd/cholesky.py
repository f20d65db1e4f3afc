"""Dense float64 Cholesky factor/solve of the reduced camera system (HIP, MFMA)."""

from __future__ import annotations

import torch

from . import _lib
from .hip_diagonal import LinAlgError


def cholesky_f64(M: torch.Tensor) -> torch.Tensor:
    """Lower Cholesky factor of a symmetric positive definite float64 matrix (lower
    triangle meaningful).  Raises LinAlgError on a non-positive pivot."""
    if M.ndim != 2 or M.shape[0] != M.shape[1] or M.dtype != torch.float64 or M.device.type != "cuda":
        raise ValueError("cholesky_f64 expects a square float64 GPU matrix")
    L = M.contiguous().clone()
    n = L.shape[0]
    status = torch.zeros(1, dtype=torch.int32, device=M.device)
    lib = _lib.load()
    _lib.check(lib.tlh_cholesky_f64(L.data_ptr(), n, n, status.data_ptr(), _lib.stream_handle(M.device)),
               "tlh_cholesky_f64")
    if int(status.item()) != 0:
        raise LinAlgError("matrix is not positive definite")
    return L


def cholesky_solve_f64(M: torch.Tensor, rhs: torch.Tensor) -> torch.Tensor:
    """Solve M x = rhs for SPD M: MFMA Cholesky factorisation, then forward and backward
    substitution over one workgroup per 64-row block.  One status read covers both
    (LinAlgError on a non-positive pivot, as torch.linalg.cholesky raises)."""
    if M.ndim != 2 or M.shape[0] != M.shape[1] or M.dtype != torch.float64 or M.device.type != "cuda":
        raise ValueError("cholesky_solve_f64 expects a square float64 GPU matrix")
    n = M.shape[0]
    if rhs.shape != (n,):
        raise ValueError(f"rhs must have shape ({n},)")
    L = M.contiguous().clone()
    x = rhs.to(torch.float64).contiguous().clone()
    status = torch.zeros(1, dtype=torch.int32, device=M.device)
    work = torch.empty(2 * max(1, n), dtype=torch.float64, device=M.device)
    lib = _lib.load()
    stream = _lib.stream_handle(M.device)
    _lib.check(lib.tlh_cholesky_f64(L.data_ptr(), n, n, status.data_ptr(), stream), "tlh_cholesky_f64")
    _lib.check(lib.tlh_cholesky_solve_blocks_f64(L.data_ptr(), n, n, x.data_ptr(), work.data_ptr(),
                                                 status.data_ptr(), stream), "tlh_cholesky_solve_blocks_f64")
    code = int(status.item())
    if code & 1:
        raise LinAlgError("matrix is not positive definite")
    if code & 2:
        # a block's bounded wait gave up (the blocks solve relies on its workgroups
        # being scheduled while earlier blocks are still running): solve again with
        # the one-workgroup substitution, which has no cross-workgroup waits
        x = rhs.to(torch.float64).contiguous().clone()
        _lib.check(lib.tlh_cholesky_solve_f64(L.data_ptr(), n, n, x.data_ptr(), stream), "tlh_cholesky_solve_f64")
    return x
