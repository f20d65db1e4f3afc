"""HIP PCG with device-side control and hipGraph-replayed iterations.

Drop-in for ``warp_pcg(matvec, rhs, *, preconditioner, config)``
(warp_pcg.py:95-214).  One iteration -- ``Ap = matvec(p)``, the curvature
``p.Ap``, the guarded step, ``z = M^-1 r``, ``r.z``, ``r.r``, the guarded
direction update and the status machine -- runs as a handful of kernels
(csrc/pcg.hip) with float64 scalars kept on the device, is captured once into a
``torch.cuda.CUDAGraph`` (a hipGraph on ROCm), and is replayed in batches of
``config.graph_batch_iterations`` between host reads of the control word.
The host-side control flow (true-residual checks, drift restarts, reasons)
follows warp_pcg._run line by line.

Fast paths: a ``HipBlockJacobi`` preconditioner is applied inside the step
kernel; any other callable is called (and captured) like in the reference.
The matvec is any graph-capturable callable returning a new tensor.
"""

from __future__ import annotations

import ctypes
import math
import threading

import torch

from . import _lib
from .hip_diagonal import HipBlockJacobi

_work_streams = threading.local()


def _work_stream(device):
    streams = getattr(_work_streams, "streams", None)
    if streams is None:
        streams = _work_streams.streams = {}
    if device not in streams:
        streams[device] = torch.cuda.Stream(device=device)
    return streams[device]


class PcgState:
    """Device control word, float64 scalars and reduction scratch of one solve."""

    _scratch: dict = {}

    def __init__(self, device):
        self.control = torch.zeros(4, dtype=torch.int32, device=device)
        self.scalars = torch.zeros(_lib.PCG_NSCALARS, dtype=torch.float64, device=device)
        self.partials = torch.zeros(2 * _lib.TLH_PCG_MAX_BLOCKS, dtype=torch.float64, device=device)
        self.counter = torch.zeros(4, dtype=torch.int32, device=device)
        self.desc = _lib.PcgStateDesc(self.control.data_ptr(), self.scalars.data_ptr(),
                                      self.partials.data_ptr(), self.counter.data_ptr())
        self.host = torch.zeros(8, dtype=torch.float64).pin_memory()

    @classmethod
    def scratch(cls, device) -> "PcgState":
        key = (threading.get_ident(), str(device))
        if key not in cls._scratch:
            cls._scratch[key] = PcgState(device)
        return cls._scratch[key]


def hip_pcg(matvec, rhs, *, preconditioner, config, use_graph: bool = True):
    """Solve A x = rhs by PCG on the GPU; returns a PCGResult like warp_pcg."""
    if rhs.device.type != "cuda" or rhs.dtype not in (torch.float32, torch.float64):
        raise ValueError("HIP PCG requires GPU float32 or float64 tensors")
    caller = torch.cuda.current_stream(rhs.device)
    stream = _work_stream(rhs.device)
    stream.wait_stream(caller)
    with torch.cuda.stream(stream):
        result = _run(matvec, rhs, preconditioner=preconditioner, config=config, use_graph=use_graph)
    caller.wait_stream(stream)
    result.solution.record_stream(caller)
    return result


def _run(matvec, rhs, *, preconditioner, config, use_graph):
    from .solvers import PCGResult

    lib = _lib.load()
    device = rhs.device
    n = rhs.numel()
    code = _lib.dtype_code(rhs.dtype)
    rhs = rhs.contiguous()
    st = PcgState(device)
    sd = ctypes.byref(st.desc)
    stream = torch.cuda.current_stream(device)
    sh = stream.cuda_stream
    fused = isinstance(preconditioner, HipBlockJacobi)
    apply = preconditioner
    pc = ctypes.byref(preconditioner.desc) if fused else None

    def dot_into(a, b, slot):
        _lib.check(lib.tlh_dot(a.data_ptr(), b.data_ptr(), n, code, sd, slot, sh), "tlh_dot")

    def read_scalars(*slots):
        st.host[: len(slots)].copy_(st.scalars[list(slots)], non_blocking=False)
        return [float(v) for v in st.host[: len(slots)].tolist()]

    x = torch.zeros_like(rhs)
    r = rhs.clone()
    dot_into(rhs, rhs, _lib.PCG_AUX)
    status = getattr(preconditioner, "status", None) if getattr(preconditioner, "deferred", False) else None
    if status is None:
        (rhs_norm2,) = read_scalars(_lib.PCG_AUX)
    else:  # the preconditioner's deferred positive-definiteness flag rides on the same read
        host = torch.stack((st.scalars[_lib.PCG_AUX], status[0].to(torch.float64))).cpu()
        rhs_norm2, bad = float(host[0]), float(host[1])
        if bad != 0:
            from .hip_diagonal import LinAlgError

            raise LinAlgError("damped block is not positive definite")
    rhs_norm = math.sqrt(rhs_norm2) if rhs_norm2 >= 0 else math.nan
    threshold = max(config.atol, config.rtol * rhs_norm)
    if not math.isfinite(rhs_norm) or rhs_norm <= threshold:
        finite = math.isfinite(rhs_norm)
        return PCGResult(x, 0, rhs_norm, finite, "residual_tolerance" if finite else "nonfinite", rhs_norm=rhs_norm)

    z = torch.empty_like(rhs)
    p = torch.empty_like(rhs)
    true_r = torch.empty_like(rhs)

    def reset(src, zero_x):
        if fused or apply is None:
            _lib.check(lib.tlh_pcg_reset(pc, src.data_ptr(), x.data_ptr() if zero_x else None, r.data_ptr(),
                                         z.data_ptr(), p.data_ptr(), n, code, sd, sh), "tlh_pcg_reset")
        else:
            if zero_x:
                x.zero_()
            r.copy_(src)
            p.copy_(apply(r))
            dot_into(r, p, _lib.PCG_RZ)

    reset(rhs, True)
    st.scalars[_lib.PCG_THR2] = threshold * threshold
    st.control.copy_(torch.tensor([0, 0, config.max_iterations, 0], dtype=torch.int32))

    small = (fused or apply is None) and n <= _lib.PCG_SMALL_MAX_N

    def iteration():
        ap = matvec(p)
        if small:  # whole iteration in one workgroup (curvature, step, M^-1, direction, status)
            _lib.check(lib.tlh_pcg_iteration_small(pc, 0 if fused else 1, x.data_ptr(), r.data_ptr(), z.data_ptr(),
                                                   p.data_ptr(), ap.data_ptr(), n, code, sd, sh),
                       "tlh_pcg_iteration_small")
            return
        dot_into(p, ap, _lib.PCG_CURVATURE)
        if fused:
            _lib.check(lib.tlh_pcg_step(pc, 0, x.data_ptr(), r.data_ptr(), z.data_ptr(), p.data_ptr(),
                                        ap.data_ptr(), n, code, sd, sh), "tlh_pcg_step")
        elif apply is None:
            _lib.check(lib.tlh_pcg_step(None, 1, x.data_ptr(), r.data_ptr(), z.data_ptr(), p.data_ptr(),
                                        ap.data_ptr(), n, code, sd, sh), "tlh_pcg_step")
        else:
            _lib.check(lib.tlh_pcg_step(None, 2, x.data_ptr(), r.data_ptr(), z.data_ptr(), p.data_ptr(),
                                        ap.data_ptr(), n, code, sd, sh), "tlh_pcg_step")
            z.copy_(apply(r))
            dot_into(r, z, _lib.PCG_NEW_RZ)
            dot_into(r, r, _lib.PCG_NORM2)
        _lib.check(lib.tlh_pcg_direction(p.data_ptr(), z.data_ptr(), n, code, sd, sh), "tlh_pcg_direction")

    # Warm up once (as warp_pcg.py:162-164), capture one iteration, reset.  The
    # capture is begun on the work stream directly: torch.cuda.graph's context
    # manager would add a device synchronize, a garbage collection and an
    # allocator cache flush to every solve.  Operators made only of this
    # library's kernels (hip_graph_safe) need no warm-up: nothing is initialised
    # lazily, and the capture records without executing, so the state set by
    # the reset above is still intact.
    owner = getattr(getattr(matvec, "func", matvec), "__self__", None)
    warm = not (use_graph and getattr(owner, "hip_graph_safe", False) and (fused or apply is None))
    if warm:
        iteration()
    graph = None
    if use_graph:
        graph = torch.cuda.CUDAGraph()
        graph.capture_begin()
        try:
            iteration()
        finally:
            graph.capture_end()
    if warm:
        reset(rhs, True)
        st.control.zero_()

    host_control = torch.zeros(4, dtype=torch.int32).pin_memory()
    iterations = checks = restarts = 0
    reason = "max_iterations"
    true_norm = math.nan
    while iterations < config.max_iterations:
        interval = config.residual_check_interval
        stop_at = min(
            config.max_iterations,
            ((iterations // interval + 1) * interval) if interval else config.max_iterations,
        )
        host_control.copy_(torch.tensor([iterations, 0, stop_at, 0], dtype=torch.int32))
        st.control.copy_(host_control, non_blocking=True)
        for _ in range(min(config.graph_batch_iterations, stop_at - iterations)):
            if graph is not None:
                graph.replay()
            else:
                iteration()
        iterations, status, _, _ = st.control.tolist()
        if status == 0 and iterations < stop_at:
            continue
        ax = matvec(x)
        _lib.check(lib.tlh_pcg_true_residual(rhs.data_ptr(), ax.data_ptr(), r.data_ptr(), true_r.data_ptr(), n,
                                             code, sd, sh), "tlh_pcg_true_residual")
        true2, drift2 = read_scalars(_lib.PCG_TRUE2, _lib.PCG_DRIFT2)
        true_norm = math.sqrt(true2) if true2 >= 0 else math.nan
        checks += 1
        if not math.isfinite(true_norm):
            reason = "nonfinite"
            break
        if true_norm <= threshold:
            return PCGResult(x, iterations, true_norm, True, "residual_tolerance", restarts, checks, rhs_norm)
        if status in (2, 3):
            reason = "nonfinite" if status == 2 else "nonpositive_curvature"
            break
        if iterations >= config.max_iterations:
            break
        drift = math.sqrt(drift2) if drift2 >= 0 else math.nan
        if (
            status == 1
            or config.restart_strategy == "periodic"
            or drift > config.residual_drift_tolerance * max(true_norm, threshold)
        ):
            reset(true_r, False)
            restarts += 1
    return PCGResult(x, iterations, true_norm, False, reason, restarts, checks, rhs_norm)
