"""Dense float64 Cholesky factor/solve of the reduced camera system (HIP, MFMA)."""

from __future__ import annotations

import torch

from . import _lib
from .hip_diagonal import LinAlgError

# status bits: 1 non-positive pivot, 2 a wait of the blocks solve gave up, 4 a wait of
# the tile factorisation gave up
_NOT_PD, _SOLVE_WAIT, _FACTOR_WAIT = 1, 2, 4


def _factor(L: torch.Tensor, status: torch.Tensor, method: str = "tiles") -> None:
    """Factor L in place on the current stream (no host read): the persistent tile
    launch, or the panel launches (method="panels", or when the device cannot keep
    the tile launch resident)."""
    n = L.shape[0]
    lib = _lib.load()
    stream = _lib.stream_handle(L.device)
    if method == "tiles":
        nbytes = int(lib.tlh_cholesky_tiles_workspace(n))
        work = torch.empty(max(1, nbytes), dtype=torch.uint8, device=L.device)
        code = lib.tlh_cholesky_tiles_f64(L.data_ptr(), n, n, work.data_ptr(), nbytes, status.data_ptr(), stream)
        if code != _lib.TLH_ELDS:
            _lib.check(code, "tlh_cholesky_tiles_f64")
            return
    elif method != "panels":
        raise ValueError("method must be tiles or panels")
    _lib.check(lib.tlh_cholesky_f64(L.data_ptr(), n, n, status.data_ptr(), stream), "tlh_cholesky_f64")


def _check_matrix(M: torch.Tensor, name: str) -> None:
    if M.ndim != 2 or M.shape[0] != M.shape[1] or M.dtype != torch.float64 or M.device.type != "cuda":
        raise ValueError(f"{name} expects a square float64 GPU matrix")


def cholesky_f64(M: torch.Tensor, *, method: str = "tiles") -> torch.Tensor:
    """Lower Cholesky factor of a symmetric positive definite float64 matrix (lower
    triangle meaningful).  Raises LinAlgError on a non-positive pivot."""
    _check_matrix(M, "cholesky_f64")
    L = M.contiguous().clone()
    status = torch.zeros(1, dtype=torch.int32, device=M.device)
    _factor(L, status, method)
    code = int(status.item())
    if code & _FACTOR_WAIT:  # the tile launch could not complete: the panel launches need no waits
        L = M.contiguous().clone()
        status.zero_()
        _factor(L, status, "panels")
        code = int(status.item())
    if code & _NOT_PD:
        raise LinAlgError("matrix is not positive definite")
    return L


def cholesky_solve_f64(M: torch.Tensor, rhs: torch.Tensor, *, method: str = "tiles") -> torch.Tensor:
    """Solve M x = rhs for SPD M: Cholesky factorisation, then forward and backward
    substitution over one workgroup per 64-row block.  One status read covers both
    (LinAlgError on a non-positive pivot, as torch.linalg.cholesky raises)."""
    _check_matrix(M, "cholesky_solve_f64")
    n = M.shape[0]
    if rhs.shape != (n,):
        raise ValueError(f"rhs must have shape ({n},)")
    lib = _lib.load()
    stream = _lib.stream_handle(M.device)
    status = torch.zeros(1, dtype=torch.int32, device=M.device)
    work = torch.empty(2 * max(1, n), dtype=torch.float64, device=M.device)

    def factor_and_solve(how):
        L = M.contiguous().clone()
        x = rhs.to(torch.float64).contiguous().clone()
        _factor(L, status, how)
        _lib.check(lib.tlh_cholesky_solve_blocks_f64(L.data_ptr(), n, n, x.data_ptr(), work.data_ptr(),
                                                     status.data_ptr(), stream), "tlh_cholesky_solve_blocks_f64")
        return L, x

    L, x = factor_and_solve(method)
    code = int(status.item())
    if code & _FACTOR_WAIT:
        status.zero_()
        L, x = factor_and_solve("panels")
        code = int(status.item())
    if code & _NOT_PD:
        raise LinAlgError("matrix is not positive definite")
    if code & _SOLVE_WAIT:
        # a block's bounded wait gave up (the blocks solve relies on its workgroups
        # being scheduled while earlier blocks are still running): solve again with
        # the one-workgroup substitution, which has no cross-workgroup waits
        x = rhs.to(torch.float64).contiguous().clone()
        _lib.check(lib.tlh_cholesky_solve_f64(L.data_ptr(), n, n, x.data_ptr(), stream), "tlh_cholesky_solve_f64")
    return x
