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
The matvec is any graph-capturable callable returning a new tensor.  When it
is one of this library's operators (an object with ``pcg_operator``: the
Schur system's ``matvec``, a normal linearization's ``normal_matvec``, bare or
as a ``functools.partial`` binding ``damping``) and the preconditioner is
fused or absent, a batch is issued natively by ``tlh_pcg_iterations`` (the
same launches as the captured iteration, in the same order, one host call per
batch) instead of graph replays.  Native and eager solves run on the caller's
stream; a captured solve runs on a per-device work stream (capture cannot use
the default stream) joined to the caller's stream by events.
"""

from __future__ import annotations

import ctypes
import functools
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
        # staging of the control word's window rewrites (host -> device, non-blocking)
        self.host_control = torch.zeros(4, dtype=torch.int32).pin_memory()

    @classmethod
    def scratch(cls, device) -> "PcgState":
        key = (threading.get_ident(), str(device))
        if key not in cls._scratch:
            cls._scratch[key] = PcgState(device)
        return cls._scratch[key]

    _solver: dict = {}

    @classmethod
    def for_solve(cls, device) -> "PcgState":
        """The state of this thread's PCG solves on a device, reused by every solve:
        tlh_pcg_begin and tlh_pcg_reset initialise it, every grid reduction leaves
        its counter at zero, and a solve reads it back before it returns."""
        key = (threading.get_ident(), str(device))
        if key not in cls._solver:
            cls._solver[key] = PcgState(device)
        return cls._solver[key]


# Replays per host read after the first batch: iterations past convergence are
# frozen by the control word but still run the operator, so later batches are short.
SMALL_BATCH = 4
# With config.stop_on_stagnation: a restart after the recurrence converged must at
# least halve the true residual of the previous such restart; otherwise the
# operator's own rounding (a float32 matvec under float64 vectors) bounds the
# attainable residual and the solve stops as "stagnated" (usable).  Without it the
# solve keeps restarting until max_iterations, as warp_pcg does.  With
# config.max_refinements, at most that many such restarts are made.
STALL_FACTOR = 0.5
# Systems up to this size run each iteration's vector work in one workgroup
# (tlh_pcg_iteration_small accepts up to _lib.PCG_SMALL_MAX_N); larger ones spread
# it over the grid, which is faster once a single compute unit's bandwidth and
# its tree reductions dominate (12,000 unknowns: 44 us in one workgroup).
SMALL_PREFERRED_N = 4096


def native_operator(matvec):
    """The ``_lib.PcgOperator`` of a library operator's matvec, or None: a bound
    method named by its owner's ``pcg_method``, bare or as a partial that binds
    only ``damping``."""
    fn, keywords = matvec, {}
    if isinstance(matvec, functools.partial):
        if matvec.args or set(matvec.keywords) - {"damping"}:
            return None
        fn, keywords = matvec.func, matvec.keywords
    owner = getattr(fn, "__self__", None)
    method = getattr(owner, "pcg_method", None)
    if method is None or getattr(fn, "__name__", None) != method:
        return None
    return owner.pcg_operator(float(keywords.get("damping", 0.0)))


def hip_pcg(matvec, rhs, *, preconditioner, config, use_graph: bool = True, first_batch: int | None = None,
            native: bool = True):
    """Solve A x = rhs by PCG on the GPU; returns a PCGResult like warp_pcg.

    ``first_batch`` (optional) is the number of iterations replayed before the
    first host read, e.g. the previous solve's count; it changes how often the
    host reads the control word, never the iterates.  ``native``: issue the
    batches of a library operator by ``tlh_pcg_iterations`` (else graph replays
    when ``use_graph``, else eager launches); all three give identical iterates.
    """
    if rhs.device.type != "cuda" or rhs.dtype not in (torch.float32, torch.float64):
        raise ValueError("HIP PCG requires GPU float32 or float64 tensors")
    fused = isinstance(preconditioner, HipBlockJacobi)
    op_desc = native_operator(matvec) if native and (fused or preconditioner is None) else None
    if op_desc is not None or not use_graph:
        # nothing is captured: the solve runs on the caller's stream, with no
        # cross-stream hand-off (two event waits, ~25 us of host time per solve)
        return _run(matvec, rhs, preconditioner=preconditioner, config=config, use_graph=False,
                    first_batch=first_batch, op_desc=op_desc)
    caller = torch.cuda.current_stream(rhs.device)
    stream = _work_stream(rhs.device)
    stream.wait_stream(caller)
    with torch.cuda.stream(stream):
        result = _run(matvec, rhs, preconditioner=preconditioner, config=config, use_graph=use_graph,
                      first_batch=first_batch, op_desc=None)
    caller.wait_stream(stream)
    result.solution.record_stream(caller)
    return result


def _run(matvec, rhs, *, preconditioner, config, use_graph, first_batch=None, op_desc=None):
    from .solvers import PCGResult

    lib = _lib.load()
    device = rhs.device
    n = rhs.numel()
    code = _lib.dtype_code(rhs.dtype)
    rhs = rhs.contiguous()
    st = PcgState.for_solve(device)
    sd = ctypes.byref(st.desc)
    stream = torch.cuda.current_stream(device)
    sh = stream.cuda_stream
    fused = isinstance(preconditioner, HipBlockJacobi)
    apply = preconditioner
    pc = ctypes.byref(preconditioner.desc) if fused else None
    interval = config.residual_check_interval

    def window_end(it):
        """Iteration at which the batch stops for a true-residual check (warp_pcg.py:177-185)."""
        return min(config.max_iterations, ((it // interval + 1) * interval) if interval else config.max_iterations)

    def dot_into(a, b, slot):
        _lib.check(lib.tlh_dot(a.data_ptr(), b.data_ptr(), n, code, sd, slot, sh), "tlh_dot")

    # a deferred positive-definiteness flag of the preconditioner is read with the first batch
    pd_flag = getattr(preconditioner, "status", None) if getattr(preconditioner, "deferred", False) else None

    def begin():
        _lib.check(lib.tlh_pcg_begin(sd, float(config.atol), float(config.rtol), window_end(0),
                                     None if pd_flag is None else pd_flag.data_ptr(), sh), "tlh_pcg_begin")

    x = torch.empty_like(rhs)  # zeroed by the reset below
    r = torch.empty_like(rhs)
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

    # threshold, early exits and the preconditioner check are decided on the device:
    # nothing is read back before the first batch of iterations
    dot_into(rhs, rhs, _lib.PCG_AUX)
    begin()
    reset(rhs, True)

    small = (fused or apply is None) and n <= SMALL_PREFERRED_N

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
    if op_desc is not None:
        use_graph = False
        ap_work = torch.empty_like(rhs)
        op_ref = ctypes.byref(op_desc)

        def run_batch(count):
            _lib.check(lib.tlh_pcg_iterations(op_ref, pc, 0 if fused else 1, x.data_ptr(), r.data_ptr(),
                                              z.data_ptr(), p.data_ptr(), ap_work.data_ptr(), n, code, sd,
                                              SMALL_PREFERRED_N, count, sh), "tlh_pcg_iterations")
    owner = getattr(getattr(matvec, "func", matvec), "__self__", None)
    warm = op_desc is None and not (use_graph and getattr(owner, "hip_graph_safe", False) and (fused or apply is None))
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
        begin()
        reset(rhs, True)

    def check_and_read():
        """True residual of the current iterate, then one host read of (rhs.rhs,
        ||rhs - Ax||^2, drift^2, iterations, status)."""
        ax = matvec(x)
        _lib.check(lib.tlh_pcg_true_residual(rhs.data_ptr(), ax.data_ptr(), r.data_ptr(), true_r.data_ptr(), n,
                                             code, sd, sh), "tlh_pcg_true_residual")
        host = st.scalars[_lib.PCG_TRUE2: _lib.PCG_STATUS + 1].cpu()  # the kernel copied the control word
        true2, drift2, rhs2, it, status = host.tolist()
        return rhs2, true2, drift2, int(it), int(status)

    host_control = st.host_control
    iterations = checks = restarts = 0
    stop_at = window_end(0)
    rhs_norm = threshold = None
    reason = "max_iterations"
    true_norm = math.nan
    converged_true = math.inf  # true residual at the last check after a converged recurrence
    refinements = 0  # restarts after a converged recurrence
    max_refinements = getattr(config, "max_refinements", None)
    batch = max(1, int(first_batch)) if first_batch else config.graph_batch_iterations
    while True:
        count = max(0, min(batch, stop_at - iterations))
        if op_desc is not None:
            if count:
                run_batch(count)
        else:
            for _ in range(count):
                if graph is not None:
                    graph.replay()
                else:
                    iteration()
        batch = min(SMALL_BATCH, config.graph_batch_iterations)
        rhs2, true2, drift2, it, status = check_and_read()
        if rhs_norm is None:
            rhs_norm = math.sqrt(rhs2) if rhs2 >= 0 else math.nan
            threshold = max(config.atol, config.rtol * rhs_norm)
            if status == 4:
                from .hip_diagonal import LinAlgError

                raise LinAlgError("damped block is not positive definite")
            if not math.isfinite(rhs_norm) or rhs_norm <= threshold:
                finite = math.isfinite(rhs_norm)
                return PCGResult(x, 0, rhs_norm, finite, "residual_tolerance" if finite else "nonfinite",
                                 rhs_norm=rhs_norm)
            if config.max_iterations <= 0:
                return PCGResult(x, 0, math.nan, False, "max_iterations", 0, 0, rhs_norm)
        iterations = it
        if status == 0 and iterations < stop_at:
            continue
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
        if status == 1:
            if getattr(config, "stop_on_stagnation", False) and true_norm > STALL_FACTOR * converged_true:
                reason = "stagnated"
                break
            if max_refinements is not None and refinements >= max_refinements:
                reason = "stagnated"
                break
            converged_true = true_norm
            refinements += 1
        drift = math.sqrt(drift2) if drift2 >= 0 else math.nan
        drifted = getattr(config, "drift_restarts", True) and drift > config.residual_drift_tolerance * max(
            true_norm, threshold)
        if status == 1 or config.restart_strategy == "periodic" or drifted:
            reset(true_r, False)
            restarts += 1
        stop_at = window_end(iterations)
        host_control.copy_(torch.tensor([iterations, 0, stop_at, 0], dtype=torch.int32))
        st.control.copy_(host_control, non_blocking=True)
    return PCGResult(x, iterations, true_norm, False, reason, restarts, checks, rhs_norm)
