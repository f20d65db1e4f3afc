"""ctypes binding of ``libtorchls_hip.so`` (C ABI: include/torchls_hip.h).

PyTorch is imported first so that the library resolves ``libamdhip64.so.7`` to
the HIP runtime PyTorch already loaded (one runtime, one set of streams).
Loading fails loudly when the library is missing: there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import torch  # noqa: F401  (must be loaded before the HIP library)

LIBRARY = Path(__file__).resolve().parent / "libtorchls_hip.so"

TLH_F32 = 0
TLH_F64 = 1
TLH_MAX_SLOTS = 3
TLH_TILE_THREADS = 256
TLH_WAVE_THREADS = 512
TLH_WAVE_TILE = 2048
PCG_SMALL_MAX_N = 16384
TLH_REDUCE_SPARSE = 0
TLH_REDUCE_DENSE = 1
TLH_EINVAL = 1000
TLH_ELDS = 1001
TLH_ENOTFOUND = 1002

_ERRORS = {
    TLH_EINVAL: "invalid argument",
    TLH_ELDS: "LDS request exceeds the device limit",
    TLH_ENOTFOUND: "no kernel for this block shape",
}

c_void_p = ctypes.c_void_p
i32 = ctypes.c_int32
i64 = ctypes.c_int64


class Slot(ctypes.Structure):
    _fields_ = [
        ("type", i32),
        ("dim", i32),
        ("mode", i32),
        ("dense_block", i32),
        ("index", c_void_p),
        ("rank", c_void_p),
        ("seg_ptr", c_void_p),
        ("seg_start", c_void_p),
        ("seg_target", c_void_p),
        ("pos_base", i64),
        ("rounds", c_void_p),
    ]


class Group(ctypes.Structure):
    _fields_ = [
        ("count", i64),
        ("rows", i32),
        ("nslots", i32),
        ("ntiles", i32),
        ("tile_cap", i32),
        ("nwg", i32),
        ("tiles_per_wg", i32),
        ("stride", i64),
        ("tile_start", c_void_p),
        ("order", c_void_p),
        ("packed", c_void_p),
        ("kernel", i32),
        ("primary", i32),
        ("out_entries", i32),
        ("stage_entries", i32),
        ("slot", Slot * TLH_MAX_SLOTS),
    ]


class Type(ctypes.Structure):
    _fields_ = [
        ("count", i64),
        ("dim", i32),
        ("dense_blocks", i32),
        ("offset", i64),
        ("sparse_partials", i64),
        ("csr_ptr", c_void_p),
        ("csr_idx", c_void_p),
        ("partials", c_void_p),
        ("direct", c_void_p),
        ("finish_list", c_void_p),
        ("finish_count", i64),
        ("csr_entries", i64),
    ]


class Operator(ctypes.Structure):
    _fields_ = [
        ("dtype", i32),
        ("ngroups", i32),
        ("ntypes", i32),
        ("items", i32),
        ("size", i64),
        ("groups", ctypes.POINTER(Group)),
        ("types", ctypes.POINTER(Type)),
    ]


TLH_MAX_TYPES = 8
TLH_PCG_MAX_BLOCKS = 1024
PCG_RZ, PCG_THR2, PCG_CURVATURE, PCG_NEW_RZ, PCG_NORM2, PCG_TRUE2, PCG_DRIFT2, PCG_AUX = range(8)
PCG_ITERATIONS, PCG_STATUS = 8, 9
PCG_NSCALARS = 10


class BlockDiag(ctypes.Structure):
    _fields_ = [
        ("ntypes", i32),
        ("dim", i32 * TLH_MAX_TYPES),
        ("count", i64 * TLH_MAX_TYPES),
        ("offset", i64 * TLH_MAX_TYPES),
        ("inv", c_void_p * TLH_MAX_TYPES),
    ]


class PcgStateDesc(ctypes.Structure):
    _fields_ = [("control", c_void_p), ("scalars", c_void_p), ("partials", c_void_p), ("counter", c_void_p)]


PCG_OP_NORMAL = 0
PCG_OP_SCHUR = 1


class PcgOperator(ctypes.Structure):
    _fields_ = [
        ("kind", i32),
        ("wgroup", i32),
        ("eslot", i32),
        ("type_end", i32),
        ("op", ctypes.POINTER(Operator)),
        ("cinv", c_void_p),
        ("damping", ctypes.c_double),
    ]


class LmTrust(ctypes.Structure):
    _fields_ = [
        ("enabled", i32),
        ("adaptive", i32),
        ("lambda_min", ctypes.c_double),
        ("lambda_max", ctypes.c_double),
        ("lambda_factor", ctypes.c_double),
        ("quality_min", ctypes.c_double),
    ]


class LmTrial(ctypes.Structure):
    _fields_ = [("cost", c_void_p), ("actual", c_void_p), ("sq", c_void_p), ("stats", c_void_p),
                ("finite", c_void_p)]


class DeviceInfo(ctypes.Structure):
    _fields_ = [("cu_count", i32), ("lds_per_block", i32), ("warp_size", i32), ("gcn_arch", i32)]


# name -> (restype, argtypes); every symbol declared in include/torchls_hip.h
SIGNATURES = {
    "tlh_device_info": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(DeviceInfo)]),
    "tlh_group_lds_bytes": (i64, [ctypes.POINTER(Operator), ctypes.c_int]),
    "tlh_pack_blocks": (
        ctypes.c_int,
        [ctypes.POINTER(Operator), ctypes.c_int, ctypes.POINTER(c_void_p), ctypes.c_int, c_void_p],
    ),
    "tlh_pack_blocks_scaled": (
        ctypes.c_int,
        [ctypes.POINTER(Operator), ctypes.c_int, ctypes.POINTER(c_void_p), ctypes.c_int, c_void_p, c_void_p],
    ),
    "tlh_normal_matvec": (
        ctypes.c_int,
        [ctypes.POINTER(Operator), c_void_p, c_void_p, ctypes.c_double, c_void_p],
    ),
    "tlh_diagonal_blocks": (
        ctypes.c_int,
        [ctypes.POINTER(Operator), ctypes.POINTER(c_void_p), i64, c_void_p],
    ),
    "tlh_block_inverse": (
        ctypes.c_int,
        [c_void_p, i64, ctypes.c_int, ctypes.c_double, c_void_p, c_void_p, c_void_p],
    ),
    "tlh_block_apply": (
        ctypes.c_int,
        [c_void_p, ctypes.c_int, c_void_p, i64, c_void_p, c_void_p, ctypes.c_int, c_void_p],
    ),
    "tlh_dot": (
        ctypes.c_int,
        [c_void_p, c_void_p, i64, ctypes.c_int, ctypes.POINTER(PcgStateDesc), ctypes.c_int, c_void_p],
    ),
    "tlh_between_linearize": (
        ctypes.c_int,
        [ctypes.c_int, c_void_p, i64, c_void_p, c_void_p, c_void_p, i64, i64, c_void_p, c_void_p, c_void_p,
         c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "tlh_pcg_begin": (
        ctypes.c_int,
        [ctypes.POINTER(PcgStateDesc), ctypes.c_double, ctypes.c_double, i32, c_void_p, c_void_p],
    ),
    "tlh_pcg_reset": (
        ctypes.c_int,
        [ctypes.POINTER(BlockDiag), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, i64, ctypes.c_int,
         ctypes.POINTER(PcgStateDesc), c_void_p],
    ),
    "tlh_pcg_step": (
        ctypes.c_int,
        [ctypes.POINTER(BlockDiag), ctypes.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, i64,
         ctypes.c_int, ctypes.POINTER(PcgStateDesc), c_void_p],
    ),
    "tlh_pcg_iteration_small": (
        ctypes.c_int,
        [ctypes.POINTER(BlockDiag), ctypes.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, i64,
         ctypes.c_int, ctypes.POINTER(PcgStateDesc), c_void_p],
    ),
    "tlh_pcg_direction": (
        ctypes.c_int,
        [c_void_p, c_void_p, i64, ctypes.c_int, ctypes.POINTER(PcgStateDesc), c_void_p],
    ),
    "tlh_pcg_true_residual": (
        ctypes.c_int,
        [c_void_p, c_void_p, c_void_p, c_void_p, i64, ctypes.c_int, ctypes.POINTER(PcgStateDesc), c_void_p],
    ),
    "tlh_pcg_iterations": (
        ctypes.c_int,
        [ctypes.POINTER(PcgOperator), ctypes.POINTER(BlockDiag), ctypes.c_int, c_void_p, c_void_p, c_void_p,
         c_void_p, c_void_p, i64, ctypes.c_int, ctypes.POINTER(PcgStateDesc), i64, i32, c_void_p],
    ),
    "tlh_rmatvec_f64": (
        ctypes.c_int,
        [ctypes.POINTER(Operator), ctypes.c_int, ctypes.POINTER(c_void_p), ctypes.c_int, c_void_p, ctypes.c_int,
         i64, c_void_p],
    ),
    "tlh_rmatvec_finish_f64": (ctypes.c_int, [ctypes.POINTER(Operator), c_void_p, c_void_p]),
    "tlh_squared_matvec_norm": (
        ctypes.c_int,
        [i64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(i32), ctypes.POINTER(i64), ctypes.POINTER(c_void_p),
         ctypes.POINTER(c_void_p), ctypes.c_int, c_void_p, ctypes.c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "tlh_block_matvec": (
        ctypes.c_int,
        [i64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(i32), ctypes.POINTER(i64), ctypes.POINTER(c_void_p),
         ctypes.POINTER(c_void_p), ctypes.c_int, c_void_p, ctypes.c_int, c_void_p, ctypes.c_int, c_void_p],
    ),
    "tlh_trial_objective": (
        ctypes.c_int,
        [c_void_p, ctypes.c_int, c_void_p, ctypes.c_int, i64, ctypes.c_int, ctypes.c_int, ctypes.c_double,
         c_void_p, ctypes.c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "tlh_schur_apply": (
        ctypes.c_int,
        [ctypes.POINTER(Operator), ctypes.c_int, ctypes.c_int, ctypes.c_int, c_void_p, c_void_p, c_void_p, c_void_p,
         c_void_p],
    ),
    "tlh_schur_precond": (
        ctypes.c_int,
        [ctypes.POINTER(Operator), ctypes.c_int, ctypes.c_int, c_void_p, c_void_p, ctypes.c_double, c_void_p, i64,
         c_void_p],
    ),
    "tlh_lm_decide": (i32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "tlh_vector_stats": (
        ctypes.c_int,
        [c_void_p, ctypes.c_int, c_void_p, ctypes.c_int, i64, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "tlh_copy_cast": (
        ctypes.c_int,
        [c_void_p, ctypes.c_int, c_void_p, i64, c_void_p, ctypes.c_int, c_void_p, i64, i64, ctypes.c_double,
         c_void_p],
    ),
    "tlh_major_rmatvec": (
        ctypes.c_int,
        [c_void_p, ctypes.c_int, c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_void_p, c_void_p, i64,
         c_void_p, c_void_p],
    ),
    "tlh_rmatvec_finish_range_f64": (
        ctypes.c_int, [ctypes.POINTER(Operator), ctypes.c_int, ctypes.c_int, c_void_p, c_void_p],
    ),
    "tlh_major_gram": (
        ctypes.c_int,
        [c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_void_p, c_void_p, i64, c_void_p, c_void_p],
    ),
    "tlh_schur_precond_major": (
        ctypes.c_int,
        [c_void_p, c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_void_p, c_void_p, c_void_p,
         c_void_p, i64, c_void_p, c_void_p, ctypes.c_double, c_void_p, c_void_p],
    ),
    "tlh_schur_dense": (
        ctypes.c_int,
        [ctypes.POINTER(Operator), ctypes.c_int, ctypes.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
         c_void_p, i64, c_void_p, i64, c_void_p, c_void_p],
    ),
    "tlh_finish_types": (
        ctypes.c_int,
        [ctypes.POINTER(Operator), ctypes.c_int, ctypes.c_int, c_void_p, ctypes.c_double, c_void_p, c_void_p],
    ),
    "tlh_group_tiles": (ctypes.c_int, [ctypes.POINTER(Operator), ctypes.c_int, c_void_p, c_void_p]),
    "tlh_cholesky_f64": (ctypes.c_int, [c_void_p, i64, i64, c_void_p, c_void_p]),
    "tlh_cholesky_tiles_workspace": (i64, [i64]),
    "tlh_set_wait_limit": (i32, [i64]),
    "tlh_cholesky_tiles_f64": (ctypes.c_int, [c_void_p, i64, i64, c_void_p, i64, c_void_p, c_void_p]),
    "tlh_cholesky_solve_f64": (ctypes.c_int, [c_void_p, i64, i64, c_void_p, c_void_p]),
    "tlh_cholesky_solve_blocks_f64": (ctypes.c_int, [c_void_p, i64, i64, c_void_p, c_void_p, c_void_p, c_void_p]),
    "tlh_cholesky_solve_blocks_f64_noalloc": (ctypes.c_int, [c_void_p, i64, i64, c_void_p, c_void_p, c_void_p]),
    "tlh_reprojection_linearize": (
        ctypes.c_int,
        [ctypes.c_int, ctypes.c_int, c_void_p, i64, c_void_p, c_void_p, c_void_p, c_void_p, i64, ctypes.c_double, c_void_p,
         c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "tlh_reprojection_trial": (
        ctypes.c_int,
        [ctypes.c_int, c_void_p, i64, c_void_p, c_void_p, c_void_p, c_void_p, i64, ctypes.c_double, c_void_p,
         c_void_p, ctypes.c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "tlh_se3_retract": (ctypes.c_int, [ctypes.c_int, c_void_p, ctypes.c_int, i64, c_void_p, i64, c_void_p, c_void_p]),
    "tlh_reprojection_hvp": (
        ctypes.c_int,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, i64, i64, ctypes.c_double, c_void_p, i64,
         c_void_p, c_void_p],
    ),
    "tlh_segment_sum_f64": (
        ctypes.c_int,
        [c_void_p, i64, ctypes.c_int, ctypes.c_int, c_void_p, c_void_p, i64, c_void_p, c_void_p],
    ),
    "tlh_reprojection_obs_vjp": (
        ctypes.c_int,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, i64, i64, ctypes.c_double, i64, c_void_p,
         c_void_p],
    ),
}

_lib = None


def load(path: Path | None = None) -> ctypes.CDLL:
    """Load the HIP library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(path) if path else LIBRARY
    if not path.exists():
        raise RuntimeError(
            f"{path} is missing: build it with `python -m torchls_amd.build` "
            "(the hip backend has no CPU fallback)"
        )
    lib = ctypes.CDLL(str(path))
    for name, (restype, argtypes) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _lib = lib
    return lib


def check(code: int, what: str) -> None:
    if code != 0:
        detail = _ERRORS.get(code, f"hipError_t {code}")
        raise RuntimeError(f"{what} failed: {detail}")


def stream_handle(device: torch.device) -> int:
    """The device's current HIP stream as an integer handle.  Called once per
    launch: torch's raw-stream accessor returns the handle without building a
    ``torch.cuda.Stream`` and re-validating the device (a few microseconds of
    host time per launch, ≈100 launches per LM iteration)."""
    if not isinstance(device, torch.device):
        device = torch.device(device)
    index = device.index
    if index is None:
        index = torch.cuda.current_device()
    return torch._C._cuda_getCurrentRawStream(index)


def dtype_code(dtype: torch.dtype) -> int:
    if dtype == torch.float32:
        return TLH_F32
    if dtype == torch.float64:
        return TLH_F64
    raise ValueError("the hip backend supports float32 and float64")


_device_info: dict = {}


def device_info(device: torch.device) -> DeviceInfo:
    index = device.index if device.index is not None else torch.cuda.current_device()
    if index not in _device_info:
        info = DeviceInfo()
        check(load().tlh_device_info(index, ctypes.byref(info)), "tlh_device_info")
        _device_info[index] = info
    return _device_info[index]
