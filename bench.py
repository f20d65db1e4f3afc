"""Benchmark of the HIP block-sparse normal-equation path on one MI355X.

Workload (default): synthetic bundle adjustment with 200 cameras, 50,000 points,
10 observations per point (500,000 reprojection residuals), Huber delta 1e-2,
LM with Schur-PCG on the 1,200-dimensional reduced camera system.  A step is
one LM iteration (linearization, diagonal blocks, Schur preconditioner,
hipGraph-replayed PCG, trial evaluation).  The mixed-precision HIP solve
(float32 blocks/matvec, float64 scalars, gradient and objective) is timed
against the all-float64 HIP solve on the same seed, and its final cost is
checked against the float64 solve.  The dominant kernel (the implicit Schur
matvec) is timed separately with HIP events for the roofline figure, and a
bounded float64 CPU-oracle sample is timed on the host cores.

Prints one JSON line on stdout; progress goes to stderr.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser(description=__doc__)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=4,
                   help="timed LM iterations (from the generator's initial point the first 4 reduce the cost by "
                        "more than float64 resolves; from the 5th on the predicted reduction is ~1e-17 of the cost "
                        "and acceptance is rounding noise)")
    p.add_argument("--warmup", type=int, default=2, help="LM iterations of the untimed warm-up solve")
    p.add_argument("--workload", default="ba_medium", choices=["ba_small", "ba_medium", "ba_large"])
    p.add_argument("--kernel-reps", type=int, default=50)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-large", action="store_true", help="skip the 5M-observation normal-matvec measurement")
    p.add_argument("--repeats", type=int, default=3, help="timed solves per precision (median reported)")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="only run a 1 GiB device copy 20 times (HBM counter calibration) and exit")
    p.add_argument("--quick", action="store_true",
                   help="skip the 5M-observation LM and pose-graph LM legs (profiling runs)")
    p.add_argument("--dense-only", action="store_true",
                   help="only time the dense Schur MFMA Cholesky (for counter profiles) and exit")
    return p.parse_args()


def build_problem(prob, device):
    from torchls_amd import HuberLoss, Problem, SE3Var
    from torchls_amd.residuals import reprojection

    P = Problem({"cams": torch.tensor(prob.cameras_init, device=device),
                 "points": torch.tensor(prob.points_init, device=device)}, manifolds={"cams": SE3Var})
    P.add_residual(reprojection, ("cams", torch.tensor(prob.cam_idx)), ("points", torch.tensor(prob.pt_idx)),
                   torch.tensor(prob.observations, device=device), loss=HuberLoss(prob.huber_delta))
    return P


def timed_solve(P, init, iters, mv_dtype):
    from torchls_amd import ConjugateGradientConfig, TerminationConfig

    P.parameters.update({k: v.clone() for k, v in init.items()})
    cg = ConjugateGradientConfig(matvec_dtype=mv_dtype)
    term = TerminationConfig(max_iterations=iters, cost_tolerance=0.0, gradient_tolerance=0.0,
                             parameter_tolerance=0.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, s = P.solve(return_summary=True, linear_solver="schur_cg", eliminate="points", conjugate_gradient=cg,
                   termination=term)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, s


def schur_matvec_timing(P, mv_dtype, reps):
    """Average time of the implicit Schur matvec (HIP events, launch stream) at lambda = 1e-3."""
    from torchls_amd.schur import HipSchurPlan

    an = P._analyze()
    lin = an.linearize(P.parameters)
    diag = tuple(d.double() for d in lin.jacobian.diagonal_blocks())
    plan = HipSchurPlan(an, "points", lin.residual.device, dtype=mv_dtype)
    system = plan.linearize(lin.jacobian, diag).damped(1e-3)
    x = torch.randn(plan.dim, dtype=mv_dtype, device=lin.residual.device)
    for _ in range(5):
        system.matvec(x)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(reps):
        system.matvec(x)
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / reps / 1e3


def normal_matvec_timing(P, mv_dtype, reps):
    from torchls_amd.hip_normal import HipNormalPlan

    an = P._analyze()
    lin = an.linearize(P.parameters)
    op = HipNormalPlan(lin.jacobian, dtype=mv_dtype).bind(lin.jacobian)
    x = torch.randn(lin.jacobian.shape[1], dtype=mv_dtype, device=lin.residual.device)
    y = torch.empty_like(x)
    for _ in range(5):
        op.normal_matvec_into(x, y, 1e-3)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(reps):
        op.normal_matvec_into(x, y, 1e-3)
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / reps / 1e3


def cpu_baseline(prob, iters):
    """float64 numpy oracle LM (Schur-PCG) on the same workload and initial point, host cores:
    the same number of LM iterations as the timed GPU solves, so its final cost is the
    float64 reference for the mixed-precision result."""
    from oracle.ba import BAProblem
    from oracle.lm import Termination, solve

    from threadpoolctl import threadpool_limits

    o = BAProblem(prob.cam_idx, prob.pt_idx, prob.observations, prob.n_cameras, prob.n_points,
                  huber_delta=prob.huber_delta)
    with threadpool_limits(limits=1):  # one host core: the reported core count
        t0 = time.perf_counter()
        _, s = solve(o, (prob.cameras_init, prob.points_init), linear_solver="schur_cg", eliminate=1,
                     term=Termination(max_iterations=iters, cost_tolerance=0, gradient_tolerance=0,
                                      parameter_tolerance=0))
        dt = time.perf_counter() - t0
    return s.iterations / dt, s


def large_normal_matvec(device, reps):
    """Normal matvec (float32 blocks) on the 2,000-camera / 500,000-point / 5M-observation
    workload: HIP-event time per launch (matvec kernel + finish kernels, launch stream)."""
    from torchls_amd import generators
    from torchls_amd.hip_normal import HipNormalPlan
    from torchls_amd.workloads import normal_matvec_bytes

    prob = generators.ba_large()
    P = build_problem(prob, device)
    lin = P._analyze().linearize(P.parameters)
    op = HipNormalPlan(lin.jacobian, dtype=torch.float32).bind(lin.jacobian)
    n = lin.jacobian.shape[1]
    x = torch.randn(n, dtype=torch.float32, device=device)
    y = torch.empty_like(x)
    t = event_time(lambda: op.normal_matvec_into(x, y, 1e-3), reps)
    nbytes = normal_matvec_bytes(prob.n_obs, 2, (6, 3), n, 4)
    return {"workload": "ba_large", "observations": prob.n_obs, "cameras": prob.n_cameras,
            "points": prob.n_points, "seed": prob.seed, "us_per_launch": round(t * 1e6, 2),
            "bytes_per_launch": nbytes, "achieved": round(nbytes / t / 1e9, 1), "unit": "GB/s",
            "peak": HBM_PEAK_GBS, "frac": round(nbytes / t / 1e9 / HBM_PEAK_GBS, 4)}


def lm_large(device, steps):
    """LM iterations/s of the mixed-precision and float64 HIP Schur-PCG at 5M observations
    (2,000 cameras, 500,000 points, 5 % outliers), where the blocks no longer fit the last-level
    cache: the same fixed-iteration protocol as the headline line, one warm-up solve each."""
    from torchls_amd import generators

    prob = generators.ba_large()
    P = build_problem(prob, device)
    init = {k: v.clone() for k, v in P.parameters.items()}
    out = {"workload": "ba_large", "cameras": prob.n_cameras, "points": prob.n_points, "observations": prob.n_obs,
           "seed": prob.seed, "steps": steps}
    res = {}
    for dt in (torch.float32, None):
        timed_solve(P, init, 1, dt)
        res[dt] = timed_solve(P, init, steps, dt)
    (t32, s32), (t64, s64) = res[torch.float32], res[None]
    out.update({"mixed_lm_iters_per_s": round(s32.iterations / t32, 3),
                "float64_lm_iters_per_s": round(s64.iterations / t64, 3),
                "mixed_over_float64": round((s32.iterations / t32) / (s64.iterations / t64), 3),
                "final_cost_mixed": s32.final_cost, "final_cost_float64": s64.final_cost,
                "final_cost_rel_err": abs(s32.final_cost - s64.final_cost) / abs(s64.final_cost),
                "pcg_iterations_mixed": s32.linear_iterations, "pcg_iterations_float64": s64.linear_iterations,
                "accepted_rejected_mixed": [s32.accepted_steps, s32.rejected_steps],
                "accepted_rejected_float64": [s64.accepted_steps, s64.rejected_steps]})
    return out


def pose_graph_lm(device, steps):
    """LM iterations/s on the 10,000-pose SE(3) pose graph (odometry + 3 % loop closures,
    6x6 aliased blocks): full-system PCG with block-Jacobi, mixed precision vs float64."""
    from torchls_amd import ConjugateGradientConfig, TerminationConfig, generators
    from torchls_amd.workloads import pose_graph_problem

    pg = generators.pose_graph()
    P = pose_graph_problem(pg, device)
    init = {k: v.clone() for k, v in P.parameters.items()}

    def run(dt, iters):
        P.parameters.update({k: v.clone() for k, v in init.items()})
        term = TerminationConfig(max_iterations=iters, cost_tolerance=0.0, gradient_tolerance=0.0,
                                 parameter_tolerance=0.0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, s = P.solve(return_summary=True, linear_solver="conjugate_gradient",
                       conjugate_gradient=ConjugateGradientConfig(matvec_dtype=dt), termination=term)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, s

    res = {}
    for dt in (torch.float32, None):
        run(dt, 1)
        res[dt] = run(dt, steps)
    (t32, s32), (t64, s64) = res[torch.float32], res[None]
    return {"workload": "pose_graph", "poses": pg.n_poses, "factors": pg.n_factors, "loop_closures": pg.n_loop,
            "seed": pg.seed, "steps": steps, "linear_solver": "conjugate_gradient (block-Jacobi)",
            "mixed_lm_iters_per_s": round(s32.iterations / t32, 3),
            "float64_lm_iters_per_s": round(s64.iterations / t64, 3),
            "mixed_over_float64": round((s32.iterations / t32) / (s64.iterations / t64), 3),
            "initial_cost": s64.initial_cost, "final_cost_mixed": s32.final_cost, "final_cost_float64": s64.final_cost,
            "final_cost_rel_err": abs(s32.final_cost - s64.final_cost) / abs(s64.final_cost),
            "pcg_iterations_mixed": s32.linear_iterations, "pcg_iterations_float64": s64.linear_iterations}


def implicit_large(device, lm_steps):
    """Differentiable solve at 1M observations (500 cameras, 100,000 points): LM with the
    full-system PCG, then the implicit gradient of sum ||log(T_true^-1 T)||^2 over the
    cameras w.r.t. every observation by one adjoint PCG solve.  Forward and backward are
    timed separately; the finite-difference check of these gradients runs at a small size
    in the GPU tests (float64 Hessian polishing does not scale to this size)."""
    from torchls_amd import ConjugateGradientConfig, HuberLoss, Problem, SE3Var, TerminationConfig, generators
    from torchls_amd.geometry import se3_log
    from torchls_amd.implicit import solve_implicit
    from torchls_amd.residuals import reprojection

    prob = generators.ba_differentiable()
    obs = torch.tensor(prob.observations, device=device, requires_grad=True)
    P = Problem({"cams": torch.tensor(prob.cameras_init, device=device),
                 "points": torch.tensor(prob.points_init, device=device)}, manifolds={"cams": SE3Var})
    P.add_residual(reprojection, ("cams", torch.tensor(prob.cam_idx)), ("points", torch.tensor(prob.pt_idx)), obs,
                   loss=HuberLoss(1e-2))
    # gauge: priors on one camera and one point
    P.add_residual(lambda c, r: (c[:3, :] - r[:3, :]).reshape(-1), ("cams", 0),
                   torch.tensor(prob.cameras_true[0], device=device))
    P.add_residual(lambda q, r: q - r, ("points", 0), torch.tensor(prob.points_true[0], device=device))
    truth = torch.tensor(prob.cameras_true, device=device)
    opts = dict(conjugate_gradient=ConjugateGradientConfig(matvec_dtype=torch.float32),
                termination=TerminationConfig(max_iterations=lm_steps, cost_tolerance=0.0, gradient_tolerance=0.0,
                                              parameter_tolerance=0.0))
    init = {k: v.clone() for k, v in P.parameters.items()}
    out = {}
    for rep in range(2):  # the first run pays the symbolic analysis
        P.parameters.update({k: v.clone() for k, v in init.items()})
        obs.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sol = solve_implicit(P, obs, **opts)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss = se3_log(torch.linalg.solve(truth, sol["cams"])).square().sum()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        out = {"forward_s": round(t1 - t0, 4), "backward_s": round(t2 - t1, 4)}
    g = obs.grad
    return {"workload": "ba_differentiable", "cameras": prob.n_cameras, "points": prob.n_points,
            "observations": prob.n_obs, "seed": prob.seed, "lm_steps": lm_steps, **out,
            "outer_loss": float(loss.detach()), "grad_norm": float(g.norm()), "grad_finite": bool(torch.isfinite(g).all())}


def dense_schur_cholesky(device, reps):
    """MFMA float64 Cholesky of the 1,536 x 1,536 reduced camera system of a 256-camera
    problem (dense Schur path): HIP-event time per factorisation and n^3/3 flop rate."""
    from torchls_amd import _lib, generators
    from torchls_amd.schur import HipSchurPlan

    prob = generators.ba_dense_schur()
    P = build_problem(prob, device)
    an = P._analyze()
    lin = an.linearize(P.parameters)
    diag = tuple(d.double() for d in lin.jacobian.diagonal_blocks())
    plan = HipSchurPlan(an, "points", device, dtype=torch.float64)
    system = plan.linearize(lin.jacobian, diag, dense=True).damped(1e-3)
    M = system.dense_matrix()
    n = M.shape[0]
    t_assembly = event_time(system.dense_matrix, reps)
    work = M.clone()
    status = torch.zeros(1, dtype=torch.int32, device=device)
    lib = _lib.load()

    def factor():
        work.copy_(M)
        _lib.check(lib.tlh_cholesky_f64(work.data_ptr(), n, n, status.data_ptr(), _lib.stream_handle(device)),
                   "tlh_cholesky_f64")

    t_all = event_time(factor, reps)
    t_copy = event_time(lambda: work.copy_(M), reps)
    t = max(t_all - t_copy, 1e-9)
    if int(status.item()) != 0:
        raise RuntimeError("dense Schur matrix is not positive definite")
    rhs = torch.randn(n, dtype=torch.float64, device=device)
    x = rhs.clone()
    swork = torch.empty(2 * n, dtype=torch.float64, device=device)

    def solve():
        x.copy_(rhs)
        _lib.check(lib.tlh_cholesky_solve_blocks_f64(work.data_ptr(), n, n, x.data_ptr(), swork.data_ptr(),
                                                     status.data_ptr(), _lib.stream_handle(device)),
                   "tlh_cholesky_solve_blocks_f64")

    t_solve = max(event_time(solve, reps) - event_time(lambda: x.copy_(rhs), reps), 1e-9)
    if int(status.item()) != 0:
        raise RuntimeError("dense Schur triangular solve did not complete")
    flops = n ** 3 / 3
    return {"workload": "ba_dense_schur", "cameras": prob.n_cameras, "n": n, "us_per_factorisation": round(t * 1e6, 1),
            "tflops": round(flops / t / 1e12, 3), "peak_tflops_fp64_mfma": 78.6,
            "frac": round(flops / t / 1e12 / 78.6, 4), "us_per_assembly": round(t_assembly * 1e6, 1),
            "us_per_solve": round(t_solve * 1e6, 1)}


TRAFFIC_SUMMARY = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "bench_hbm_traffic.md")


def profiled_traffic(kernels):
    """HBM bytes per launch of the given (kernel name prefix, grid) launches, summed, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this benchmark (copy-calibrated);
    None when the summary or a kernel is missing."""
    try:
        lines = open(TRAFFIC_SUMMARY).read().splitlines()
    except OSError:
        return None
    rows = []
    for line in lines:
        cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and cells[1].isdigit() and cells[2].isdigit():
            rows.append((cells[0], int(cells[1]), int(cells[2]), float(cells[5])))
    total = 0.0
    for name, grid in kernels:
        match = [r for r in rows if r[0] == name and (grid is None or r[1] == grid)]
        if not match:
            return None
        total += max(match, key=lambda r: r[2])[3] * 1e6  # the most frequent launch of that kernel
    return total


def event_time(fn, reps):
    """Average seconds per call: HIP events on the current stream around ``reps`` calls."""
    for _ in range(5):
        fn()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(reps):
        fn()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / reps / 1e3


def calibrate_copy():
    """A plain device copy of known size, for scaling the HBM traffic counters."""
    n = 256 * 1024 * 1024
    src = torch.randn(n, device="cuda:0")
    dst = torch.empty_like(src)
    t = event_time(lambda: dst.copy_(src), 20)
    log(f"copy of {4 * n} bytes: {t * 1e6:.1f} us, {2 * 4 * n / t / 1e9:.0f} GB/s (read + write)")


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.dense_only:
        print(json.dumps(dense_schur_cholesky(torch.device("cuda:0"), 20)), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the sharding design is "
              "described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    from torchls_amd import generators
    from torchls_amd.workloads import normal_matvec_bytes

    device = torch.device("cuda:0")
    torch.manual_seed(0)
    prob = getattr(generators, args.workload)()
    log(f"workload {args.workload}: seed={prob.seed} cams={prob.n_cameras} points={prob.n_points} obs={prob.n_obs}")
    P = build_problem(prob, device)
    init = {k: v.clone() for k, v in P.parameters.items()}

    # warm-up solves: symbolic plans (cached on the problem's analysis and reused
    # by later solves, like a serving loop), kernel loading, graph capture paths
    cold = {}
    for dt in (torch.float32, None):
        cold[dt] = timed_solve(P, init, max(1, args.warmup), dt)[0]
    warm = timed_solve(P, init, max(1, args.warmup), torch.float32)[0]
    plan_s = cold[torch.float32] - warm
    log(f"warm-up done: first mixed solve {cold[torch.float32]:.3f}s, warm {warm:.3f}s "
        f"(symbolic analysis and first-call cost {plan_s:.3f}s)")
    runs = {torch.float32: [], None: []}
    for _ in range(max(1, args.repeats)):
        for dt in (torch.float32, None):
            runs[dt].append(timed_solve(P, init, args.steps, dt))
    t_mixed, s_mixed = sorted(runs[torch.float32], key=lambda r: r[0])[len(runs[torch.float32]) // 2]
    t_f64, s_f64 = sorted(runs[None], key=lambda r: r[0])[len(runs[None]) // 2]
    log("timed solves (s): mixed " + " ".join(f"{r[0]:.4f}" for r in runs[torch.float32]) + "; float64 " +
        " ".join(f"{r[0]:.4f}" for r in runs[None]))
    log(f"mixed: {args.steps} LM iterations in {t_mixed:.4f}s (median of {len(runs[None])}) "
        f"final_cost={s_mixed.final_cost:.12e}")
    log(f"float64: {args.steps} LM iterations in {t_f64:.4f}s final_cost={s_f64.final_cost:.12e}")
    rel = abs(s_mixed.final_cost - s_f64.final_cost) / abs(s_f64.final_cost)
    iters_mixed = s_mixed.iterations
    iters_f64 = s_f64.iterations
    lm_mixed = iters_mixed / t_mixed
    lm_f64 = iters_f64 / t_f64

    P.parameters.update({k: v.clone() for k, v in init.items()})
    t_schur = schur_matvec_timing(P, torch.float32, args.kernel_reps)
    n_c, n_p, N = prob.n_cameras, prob.n_points, prob.n_obs
    # implicit Schur matvec algorithmic bytes: blocks + two int32 ids per observation,
    # C^-1 (float64 upper 3x3) per point, camera vector in/out
    schur_bytes = N * (2 * 9 * 4 + 8) + n_p * 6 * 8 + 2 * 6 * n_c * 4
    # the camera finish of the Schur matvec: 8 columns per 256-thread workgroup
    schur_traffic = profiled_traffic([("void schur_tile_kernel<float, 2, 6, 3, true, 1, 0, 1024, true>", None),
                                      ("void finish_jobs_kernel<float>", -(-n_c * 6 // 8) * 256)])
    schur_gbs = schur_bytes / t_schur / 1e9
    t_normal = normal_matvec_timing(P, torch.float32, args.kernel_reps)
    n = 6 * n_c + 3 * n_p
    normal_gbs = normal_matvec_bytes(N, 2, (6, 3), n, 4) / t_normal / 1e9
    log(f"schur matvec {t_schur * 1e6:.2f} us ({schur_gbs:.1f} GB/s); normal matvec {t_normal * 1e6:.2f} us "
        f"({normal_gbs:.1f} GB/s)")

    large = None
    if not args.no_large:
        large = large_normal_matvec(device, args.kernel_reps)
        log(f"normal matvec at 5M observations: {large['us_per_launch']} us ({large['achieved']} GB/s, "
            f"{100 * large['frac']:.1f}% of {HBM_PEAK_GBS:.0f} GB/s)")

    lm5m = None
    if not args.no_large and not args.quick:
        lm5m = lm_large(device, 3)
        log(f"LM at 5M observations: mixed {lm5m['mixed_lm_iters_per_s']} it/s, float64 "
            f"{lm5m['float64_lm_iters_per_s']} it/s (x{lm5m['mixed_over_float64']}), final-cost rel err "
            f"{lm5m['final_cost_rel_err']:.2e}")

    pgraph = None if args.quick else pose_graph_lm(device, 5)
    if pgraph:
        log(f"pose graph ({pgraph['poses']} poses, {pgraph['factors']} factors): mixed {pgraph['mixed_lm_iters_per_s']} "
            f"it/s, float64 {pgraph['float64_lm_iters_per_s']} it/s, final-cost rel err "
            f"{pgraph['final_cost_rel_err']:.2e}")

    chol = dense_schur_cholesky(device, 20)
    log(f"dense Schur Cholesky n={chol['n']}: {chol['us_per_factorisation']} us, {chol['tflops']} TFLOP/s")

    implicit = None if args.quick else implicit_large(device, 10)
    if implicit:
        log(f"implicit gradient at {implicit['observations']} observations: forward {implicit['forward_s']} s, "
            f"backward (adjoint) {implicit['backward_s']} s, |grad| {implicit['grad_norm']:.3e}")

    cpu = None
    oracle_rel = None
    if not args.no_cpu:
        rate, ref = cpu_baseline(prob, args.steps)
        oracle_rel = abs(s_mixed.final_cost - ref.final_cost) / abs(ref.final_cost)
        log(f"float64 CPU oracle: {ref.iterations} LM iterations at {rate:.4f} it/s, final_cost="
            f"{ref.final_cost:.12e}; mixed HIP rel err {oracle_rel:.2e}")
        cpu = {"value": round(rate, 4), "unit": "LM iterations/s", "cores": 1, "kind": "port",
               "sample": f"{ref.iterations} LM iterations (Schur-PCG, float64 numpy oracle, BLAS limited to one "
                         "thread) of the same 500,000-observation workload from the same initial point"}
    print(json.dumps({
        "metric": "LM iterations/s (mixed-precision HIP Schur-PCG), with matvec HBM GB/s",
        "value": round(lm_mixed, 3), "unit": "LM iterations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * t_mixed / max(1, iters_mixed), 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 blocks/matvec, f64 scalars",
        "data": "synthetic (seeded generator)",
        "config": {"workload": args.workload, "cameras": n_c, "points": n_p, "observations": N, "seed": prob.seed,
                   "linear_solver": "schur_cg", "huber_delta": prob.huber_delta},
        "float64_lm_iters_per_s": round(lm_f64, 3), "mixed_over_float64": round(lm_mixed / lm_f64, 3),
        "final_cost_mixed": s_mixed.final_cost, "final_cost_float64": s_f64.final_cost,
        "final_cost_rel_err": rel, "final_cost_rel_err_vs_cpu_oracle": oracle_rel,
        "accepted_rejected_mixed": [s_mixed.accepted_steps, s_mixed.rejected_steps],
        "accepted_rejected_float64": [s_f64.accepted_steps, s_f64.rejected_steps],
        "pcg_iterations_mixed": s_mixed.linear_iterations,
        "pcg_iterations_float64": s_f64.linear_iterations,
        "normal_matvec_gbs": round(normal_gbs, 1), "normal_matvec_us": round(t_normal * 1e6, 2),
        "first_solve_extra_s": round(plan_s, 4),
        "roofline": {"bound": "hbm", "achieved": round(schur_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(schur_gbs / HBM_PEAK_GBS, 4), "traffic": schur_traffic,
                     "traffic_source": "profiles/bench_hbm_traffic.md (rocprofv3 FETCH_SIZE, WRITE_SIZE)",
                     "kernel": "implicit Schur matvec (schur_tile_kernel + finish)",
                     "us_per_launch": round(t_schur * 1e6, 2), "bytes_per_launch": schur_bytes},
        "normal_matvec_large": large, "lm_large": lm5m, "pose_graph": pgraph,
        "dense_schur_cholesky": chol, "implicit_gradient": implicit,
        "cpu_baseline": cpu,
    }), flush=True)


if __name__ == "__main__":
    main()
