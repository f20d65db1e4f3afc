"""Benchmark of the HIP block-sparse normal-equation path on one MI355X.

Headline workload: synthetic bundle adjustment with 200 cameras, 50,000 points,
10 observations per point (500,000 reprojection residuals), Huber delta 1e-2,
LM with Schur-PCG on the 1,200-dimensional reduced camera system.  A step is
one LM iteration (linearization, gradient, diagonal blocks, Schur
preconditioner, hipGraph-replayed PCG, back-substitution, retraction, trial
evaluation).  From the generator's initial point the first four iterations are
productive (after them the predicted reduction is ~1e-17 of the cost, below
float64 resolution), so the K timed steps are ceil(K / 4) solves of four
iterations each, restarted from the initial point, with inputs resident in
HBM.  The mixed-precision HIP solve (float32 blocks/matvec/PCG vectors, float64
scalars, gradient and objective) is timed against the all-float64 HIP solve on
the same seed and protocol; iterations, trials and termination reasons of both
are reported, with the final-cost error against the float64 HIP solve and the
float64 CPU oracle.  The roofline figure is the normal matvec at 5,012,836
observations (2,000 cameras, the HBM-resident configuration), timed with HIP
events on its launch stream; a bounded float64 CPU-oracle sample runs on the
host cores.

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
    p.add_argument("--steps", type=int, default=20, help="timed LM iterations (in solves of --segment iterations)")
    p.add_argument("--segment", type=int, default=4,
                   help="LM iterations per timed solve from the initial point (the productive ones: from the 5th "
                        "on the predicted reduction is ~1e-17 of the cost and acceptance is rounding noise)")
    p.add_argument("--warmup", type=int, default=5, help="LM iterations of the untimed warm-up solves")
    p.add_argument("--workload", default="ba_medium", choices=["ba_small", "ba_medium", "ba_large"])
    p.add_argument("--kernel-reps", type=int, default=50)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--matvec-only", action="store_true",
                   help="only time the 5M-observation normal matvec (for counter profiles) and exit")
    p.add_argument("--lm-large-only", action="store_true",
                   help="only run the 5M-observation LM leg (for kernel-trace profiles) and exit")
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


# Both precisions run the same solver options.  A float32 operator's true residual has a
# rounding floor, usually above rtol 1e-6: max_refinements=0 ends the solve at the first
# converged recurrence that misses rtol (no restart; the float64 LM iteration corrects the
# step), and stop_on_stagnation ends restarts that stop reducing the true residual.  Float64
# solves nearly always meet rtol at that first check.  drift_restarts=False: the same floor
# makes the drift test fire at every check of an ill-conditioned float32 solve (the
# 10,000-pose graph: 85 restarts, 5,657 instead of 3,787 iterations, final cost 1.2e-6 off
# the float64 solve; without them 3,787 iterations and 4.6e-9); the bundle-adjustment
# solves and every float64 solve here restart on no drift either way.
CG_OPTIONS = dict(stop_on_stagnation=True, max_refinements=0, drift_restarts=False)


def timed_solve(P, init, iters, mv_dtype):
    from torchls_amd import ConjugateGradientConfig, TerminationConfig

    P.parameters.update({k: v.clone() for k, v in init.items()})
    cg = ConjugateGradientConfig(matvec_dtype=mv_dtype, **CG_OPTIONS)
    term = TerminationConfig(max_iterations=iters, cost_tolerance=0.0, gradient_tolerance=0.0,
                             parameter_tolerance=0.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, s = P.solve(return_summary=True, linear_solver="schur_cg", eliminate="points", conjugate_gradient=cg,
                   termination=term)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, s


def segmented_run(P, init, steps, segment, mv_dtype):
    """``steps`` LM iterations as solves of ``segment`` iterations from ``init`` (each
    solve bracketed by device synchronizations, the parameter reset outside the
    clock): (seconds, per-solve summaries)."""
    total, sums, left = 0.0, [], steps
    while left > 0:
        t, s = timed_solve(P, init, min(segment, left), mv_dtype)
        total += t
        sums.append(s)
        left -= min(segment, left)
    return total, sums


def run_stats(t, sums):
    from collections import Counter

    it = sum(s.iterations for s in sums)
    acc = sum(s.accepted_steps for s in sums)
    rej = sum(s.rejected_steps for s in sums)
    return {"seconds": round(t, 5), "iterations": it, "accepted": acc, "rejected": rej, "trials": acc + rej,
            "lm_iters_per_s": round(it / t, 3), "trials_per_s": round((acc + rej) / t, 3),
            "pcg_iterations": sum(s.linear_iterations for s in sums),
            "termination": dict(Counter(s.termination_reason for s in sums)),
            "final_cost": sums[-1].final_cost}


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


def host_cores():
    """Host threads this process may use (the scheduler affinity, capped by OMP_NUM_THREADS
    where the environment sets the machine's share)."""
    n = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS")
    return min(n, int(cap)) if cap and cap.isdigit() and int(cap) > 0 else n


def cpu_baseline(prob, iters, linear_solver="schur_cg"):
    """float64 numpy oracle LM on the same workload and initial point, BLAS on every host
    core this process may use: its final cost is the float64 reference for the HIP result."""
    from oracle.ba import BAProblem
    from oracle.lm import Termination, solve

    from threadpoolctl import threadpool_limits

    cores = host_cores()
    o = BAProblem(prob.cam_idx, prob.pt_idx, prob.observations, prob.n_cameras, prob.n_points,
                  huber_delta=prob.huber_delta)
    with threadpool_limits(limits=cores):
        t0 = time.perf_counter()
        _, s = solve(o, (prob.cameras_init, prob.points_init), linear_solver=linear_solver,
                     eliminate=1 if linear_solver.startswith("schur") else None,
                     term=Termination(max_iterations=iters, cost_tolerance=0, gradient_tolerance=0,
                                      parameter_tolerance=0))
        dt = time.perf_counter() - t0
    return s.iterations / dt, s, cores


def small_cpu_vs_hip(device, iters=20):
    """The reference's CPU-runnable configuration: 16 cameras, 500 points, 4,000 residuals,
    float64, LM with full-system PCG + block Jacobi, 20 iterations -- the float64 CPU oracle
    (host cores) against the float64 HIP solve (median of 3 after a warm-up solve)."""
    from torchls_amd import ConjugateGradientConfig, TerminationConfig, generators

    prob = generators.ba_small()
    P = build_problem(prob, device)
    init = {k: v.clone() for k, v in P.parameters.items()}
    term = TerminationConfig(max_iterations=iters, cost_tolerance=0.0, gradient_tolerance=0.0,
                             parameter_tolerance=0.0)
    times = []
    for rep in range(4):
        P.parameters.update({k: v.clone() for k, v in init.items()})
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, s = P.solve(return_summary=True, termination=term, conjugate_gradient=ConjugateGradientConfig())
        torch.cuda.synchronize()
        if rep:
            times.append(time.perf_counter() - t0)
    t = sorted(times)[1]
    rate, ref, cores = cpu_baseline(prob, iters, "conjugate_gradient")
    return {"workload": "ba_small", "cameras": prob.n_cameras, "points": prob.n_points,
            "observations": prob.n_obs, "seed": prob.seed, "iterations": s.iterations,
            "accepted_rejected": [s.accepted_steps, s.rejected_steps], "termination": s.termination_reason,
            "hip_float64_lm_iters_per_s": round(s.iterations / t, 2),
            "cpu_oracle_lm_iters_per_s": round(rate, 2), "cpu_cores": cores,
            "hip_over_cpu": round((s.iterations / t) / rate, 2),
            "final_cost_hip": s.final_cost, "final_cost_cpu_oracle": ref.final_cost,
            "final_cost_rel_err": abs(s.final_cost - ref.final_cost) / abs(ref.final_cost),
            "pcg_iterations_equal": s.linear_iterations_history == ref.linear_iterations_history}


def large_normal_matvec(device, reps):
    """Normal matvec (float32 blocks) on the 2,000-camera / 500,000-point / 5M-observation
    workload: HIP-event time per launch (matvec kernel + finish kernels) on the stream the
    kernels are launched on (the current stream)."""
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
    gbs = nbytes / t / 1e9
    wave = op.plan.plan.groups[0]
    traffic = profiled_traffic(MATVEC_TRAFFIC, [("void normal_wave_kernel<float, 2, 6, 3, 1, 4, 512, true, true, true, true>", None),
                                                ("void finish_jobs_kernel<float>", None)])
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": "profiles/normal_matvec_5M_hbm_traffic.md (rocprofv3 FETCH_SIZE, WRITE_SIZE)",
            "kernel": "normal matvec J^T J x + lambda x, float32 (normal_wave_kernel + finish_jobs_kernel)",
            "workload": "ba_large", "observations": prob.n_obs, "cameras": prob.n_cameras, "points": prob.n_points,
            "seed": prob.seed, "us_per_launch": round(t * 1e6, 2), "bytes_per_launch": nbytes,
            "bytes_per_unit": "80 B per observation (2x6 + 2x3 float32 blocks, two int32 ids) + 8 B per "
                              "tangent entry (x read, y written)", "workgroups": int(wave.nwg)}


def lm_large(device, steps):
    """LM iterations/s of the mixed-precision and float64 HIP Schur-PCG at 5M observations
    (2,000 cameras, 500,000 points, 5 % outliers), where the blocks no longer fit the last-level
    cache: ``steps`` productive iterations from the initial point, one warm-up solve each."""
    from torchls_amd import generators

    prob = generators.ba_large()
    P = build_problem(prob, device)
    init = {k: v.clone() for k, v in P.parameters.items()}
    out = {"workload": "ba_large", "cameras": prob.n_cameras, "points": prob.n_points, "observations": prob.n_obs,
           "seed": prob.seed, "steps": steps}
    res = {}
    for dt in (torch.float32, None):
        timed_solve(P, init, 1, dt)
        t, s = timed_solve(P, init, steps, dt)
        res[dt] = run_stats(t, [s])
    m, f = res[torch.float32], res[None]
    out.update({"mixed": m, "float64": f, "mixed_over_float64": round(m["lm_iters_per_s"] / f["lm_iters_per_s"], 3),
                "final_cost_rel_err": abs(m["final_cost"] - f["final_cost"]) / abs(f["final_cost"])})
    return out


def pose_graph_lm(device, steps):
    """LM on the 10,000-pose SE(3) pose graph (odometry + 3 % loop closures, 6x6 aliased
    blocks, started at the true poses perturbed by 0.01 per tangent coordinate): full-system
    PCG with block Jacobi, mixed precision vs float64.  The first three LM iterations are
    the ones whose PCG solves meet rtol within 3,000 iterations (the chain's Laplacian has a
    condition number of order n^2), so the mixed result is compared with float64 there."""
    from torchls_amd import ConjugateGradientConfig, TerminationConfig, generators
    from torchls_amd.workloads import pose_graph_problem

    pg = generators.pose_graph(init_noise=0.01)
    P = pose_graph_problem(pg, device)
    init = {k: v.clone() for k, v in P.parameters.items()}

    def run(dt, iters):
        P.parameters.update({k: v.clone() for k, v in init.items()})
        term = TerminationConfig(max_iterations=iters, cost_tolerance=0.0, gradient_tolerance=0.0,
                                 parameter_tolerance=0.0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, s = P.solve(return_summary=True, linear_solver="conjugate_gradient",
                       conjugate_gradient=ConjugateGradientConfig(matvec_dtype=dt, max_iterations=3000,
                                                                  **CG_OPTIONS),
                       termination=term)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, s

    res = {}
    for dt in (torch.float32, None):
        run(dt, 1)
        t, s = run(dt, steps)
        res[dt] = (run_stats(t, [s]), s)
    (m, s32), (f, s64) = res[torch.float32], res[None]
    return {"workload": "pose_graph", "poses": pg.n_poses, "factors": pg.n_factors, "loop_closures": pg.n_loop,
            "seed": pg.seed, "init": "true poses x Exp(N(0, 0.01^2 I))", "steps": steps,
            "linear_solver": "conjugate_gradient (block-Jacobi, max 3,000 iterations)", "mixed": m, "float64": f,
            "mixed_over_float64": round(m["lm_iters_per_s"] / f["lm_iters_per_s"], 3),
            "initial_cost": s64.initial_cost,
            "final_cost_rel_err": abs(s32.final_cost - s64.final_cost) / abs(s64.final_cost),
            "pcg_history_mixed": s32.linear_iterations_history, "pcg_history_float64": s64.linear_iterations_history,
            "pcg_converged_float64": all(s64.linear_converged_history)}


def implicit_large(device, lm_steps, probes=8):
    """Differentiable solve at 1M observations (500 cameras, 100,000 points): mixed-precision
    LM (full-system PCG), two exact-Hessian Newton polishing steps, then the implicit gradient
    of sum ||log(T_true^-1 T)||^2 over the cameras w.r.t. every observation (Gauss-Newton
    adjoint PCG refined by conjugate gradients on the exact Hessian, HIP Hessian-vector
    products).  Forward and backward are timed separately (second run; the first pays the
    symbolic analysis), and the gradient is checked against float64 central differences
    (h = 1e-6) on ``probes`` random observation coordinates."""
    import numpy as np

    from torchls_amd import ConjugateGradientConfig, HuberLoss, Problem, SE3Var, TerminationConfig, generators
    from torchls_amd.geometry import se3_log
    from torchls_amd.implicit import newton_polish, solve_implicit
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

    def loss_fn(cams):
        return se3_log(torch.linalg.solve(truth, cams)).square().sum()

    # the adjoint's Gauss-Newton solves keep their float32 refinement restarts (no
    # max_refinements): the exact-Hessian refinement starts from their solution
    opts = dict(conjugate_gradient=ConjugateGradientConfig(matvec_dtype=torch.float32, stop_on_stagnation=True),
                termination=TerminationConfig(max_iterations=lm_steps))
    init = {k: v.clone() for k, v in P.parameters.items()}
    out = {}
    for rep in range(2):  # the first run pays the symbolic analysis
        P.parameters.update({k: v.clone() for k, v in init.items()})
        obs.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sol = solve_implicit(P, obs, refine=30, polish=2, **opts)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss = loss_fn(sol["cams"])
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        out = {"forward_s": round(t1 - t0, 4), "backward_s": round(t2 - t1, 4)}
    g = obs.grad
    theta = {k: v.detach().clone() for k, v in P.parameters.items()}
    rng = np.random.default_rng(0)
    h, pairs = 1e-6, []
    for i in rng.choice(prob.n_obs, size=probes, replace=False):
        c = int(rng.integers(0, 2))
        vals = []
        for sgn in (1.0, -1.0):
            with torch.no_grad():
                obs[i, c] += sgn * h
            P.parameters.update({k: v.clone() for k, v in theta.items()})
            vals.append(float(loss_fn(newton_polish(P, steps=2, rtol=1e-10)["cams"])))
            with torch.no_grad():
                obs[i, c] -= sgn * h
        pairs.append((float(g[i, c]), (vals[0] - vals[1]) / (2 * h)))
    a = np.array(pairs)
    fd_err = float(np.linalg.norm(a[:, 0] - a[:, 1]) / np.linalg.norm(a[:, 1]))
    return {"workload": "ba_differentiable", "cameras": prob.n_cameras, "points": prob.n_points,
            "observations": prob.n_obs, "seed": prob.seed, "lm_steps": lm_steps, "polish_newton_steps": 2,
            "adjoint": "Gauss-Newton PCG + 30 exact-Hessian CG steps (HIP Hessian-vector products)", **out,
            "outer_loss": float(loss.detach()), "grad_norm": float(g.norm()),
            "grad_finite": bool(torch.isfinite(g).all()), "fd_probes": probes, "fd_step": h,
            "fd_rel_err": fd_err}

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
    tile_bytes = int(lib.tlh_cholesky_tiles_workspace(n))
    tile_work = torch.empty(tile_bytes, dtype=torch.uint8, device=device)

    def factor():  # the persistent tile launch (the dense Schur path's factorisation)
        work.copy_(M)
        _lib.check(lib.tlh_cholesky_tiles_f64(work.data_ptr(), n, n, tile_work.data_ptr(), tile_bytes,
                                              status.data_ptr(), _lib.stream_handle(device)), "tlh_cholesky_tiles_f64")

    def factor_panels():  # the panel launches (fallback)
        work.copy_(M)
        _lib.check(lib.tlh_cholesky_f64(work.data_ptr(), n, n, status.data_ptr(), _lib.stream_handle(device)),
                   "tlh_cholesky_f64")

    t_copy = event_time(lambda: work.copy_(M), reps)
    t_panels = max(event_time(factor_panels, reps) - t_copy, 1e-9)
    t = max(event_time(factor, reps) - t_copy, 1e-9)
    if int(status.item()) != 0:
        raise RuntimeError(f"dense Schur factorisation status {int(status.item())}")
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
            "kernel": "chol_tiles_kernel (one persistent cooperative launch over 64x64 tiles)",
            "us_per_factorisation_panels": round(t_panels * 1e6, 1),
            "tflops": round(flops / t / 1e12, 3), "peak_tflops_fp64_mfma": 78.6,
            "frac": round(flops / t / 1e12 / 78.6, 4), "us_per_assembly": round(t_assembly * 1e6, 1),
            "us_per_solve": round(t_solve * 1e6, 1)}


PROFILES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")
MATVEC_TRAFFIC = os.path.join(PROFILES, "normal_matvec_5M_hbm_traffic.md")


def profiled_traffic(summary, kernels):
    """HBM bytes per launch of the given (kernel name, grid or None) launches, summed, from a
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated); None when the
    summary or a kernel is missing."""
    try:
        lines = open(summary).read().splitlines()
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
    if args.matvec_only:
        print(json.dumps(large_normal_matvec(torch.device("cuda:0"), args.kernel_reps)), flush=True)
        return
    if args.lm_large_only:
        print(json.dumps(lm_large(torch.device("cuda:0"), 3)), flush=True)
        return
    from torchls_amd import generators

    device = torch.device("cuda:0")
    torch.manual_seed(0)
    prob = getattr(generators, args.workload)()
    log(f"workload {args.workload}: seed={prob.seed} cams={prob.n_cameras} points={prob.n_points} obs={prob.n_obs}")
    P = build_problem(prob, device)
    init = {k: v.clone() for k, v in P.parameters.items()}

    # warm-up solves: symbolic plans (cached on the problem's analysis and reused by
    # later solves, like a serving loop), kernel loading, graph capture paths
    cold = {}
    for dt in (torch.float32, None):
        cold[dt] = segmented_run(P, init, max(1, args.warmup), args.segment, dt)[0]
    warm = segmented_run(P, init, max(1, args.warmup), args.segment, torch.float32)[0]
    plan_s = cold[torch.float32] - warm
    log(f"warm-up done: first mixed run {cold[torch.float32]:.3f}s, warm {warm:.3f}s "
        f"(symbolic analysis and first-call cost {plan_s:.3f}s)")
    runs = {torch.float32: [], None: []}
    for _ in range(max(1, args.repeats)):
        for dt in (torch.float32, None):
            runs[dt].append(segmented_run(P, init, args.steps, args.segment, dt))
    med = {dt: sorted(r, key=lambda x: x[0])[len(r) // 2] for dt, r in runs.items()}
    mixed, f64 = run_stats(*med[torch.float32]), run_stats(*med[None])
    log("timed runs (s): mixed " + " ".join(f"{r[0]:.4f}" for r in runs[torch.float32]) + "; float64 " +
        " ".join(f"{r[0]:.4f}" for r in runs[None]))
    for name, st in (("mixed", mixed), ("float64", f64)):
        log(f"{name}: {st['iterations']} LM iterations ({st['accepted']} accepted, {st['rejected']} rejected, "
            f"{st['termination']}) in {st['seconds']:.4f}s: {st['lm_iters_per_s']} it/s, "
            f"{st['trials_per_s']} trials/s, final_cost={st['final_cost']:.12e}")
    rel = abs(mixed["final_cost"] - f64["final_cost"]) / abs(f64["final_cost"])

    P.parameters.update({k: v.clone() for k, v in init.items()})
    t_schur = schur_matvec_timing(P, torch.float32, args.kernel_reps)
    n_c, n_p, N = prob.n_cameras, prob.n_points, prob.n_obs
    # implicit Schur matvec algorithmic bytes: blocks + two int32 ids per observation,
    # C^-1 (float64 upper 3x3) per point, camera vector in/out
    schur_bytes = N * (2 * 9 * 4 + 8) + n_p * 6 * 8 + 2 * 6 * n_c * 4
    schur_gbs = schur_bytes / t_schur / 1e9
    log(f"schur matvec {t_schur * 1e6:.2f} us ({schur_gbs:.1f} GB/s effective, MALL-resident)")

    roofline = large_normal_matvec(device, args.kernel_reps)
    log(f"normal matvec at 5M observations: {roofline['us_per_launch']} us ({roofline['achieved']} GB/s, "
        f"{100 * roofline['frac']:.1f}% of {HBM_PEAK_GBS:.0f} GB/s), traffic {roofline['traffic']}")

    lm5m = None
    if not args.quick:
        lm5m = lm_large(device, 3)
        log(f"LM at 5M observations: mixed {lm5m['mixed']['lm_iters_per_s']} it/s, float64 "
            f"{lm5m['float64']['lm_iters_per_s']} it/s (x{lm5m['mixed_over_float64']}), final-cost rel err "
            f"{lm5m['final_cost_rel_err']:.2e}")

    pgraph = None if args.quick else pose_graph_lm(device, 3)
    if pgraph:
        log(f"pose graph ({pgraph['poses']} poses, {pgraph['factors']} factors): mixed "
            f"{pgraph['mixed']['lm_iters_per_s']} it/s, float64 {pgraph['float64']['lm_iters_per_s']} it/s, "
            f"final-cost rel err {pgraph['final_cost_rel_err']:.2e}")

    chol = dense_schur_cholesky(device, 20)
    log(f"dense Schur Cholesky n={chol['n']}: {chol['us_per_factorisation']} us, {chol['tflops']} TFLOP/s")

    implicit = None if args.quick else implicit_large(device, 10)
    if implicit:
        log(f"implicit gradient at {implicit['observations']} observations: forward {implicit['forward_s']} s, "
            f"backward (adjoint) {implicit['backward_s']} s, |grad| {implicit['grad_norm']:.3e}, FD rel err "
            f"{implicit['fd_rel_err']:.2e} ({implicit['fd_probes']} probes)")

    small = None
    cpu = None
    oracle_rel = None
    if not args.no_cpu:
        small = small_cpu_vs_hip(device)
        log(f"16 cameras / 4,000 residuals, 20 LM iterations: HIP float64 {small['hip_float64_lm_iters_per_s']} it/s, "
            f"CPU oracle {small['cpu_oracle_lm_iters_per_s']} it/s on {small['cpu_cores']} cores, final-cost rel err "
            f"{small['final_cost_rel_err']:.2e}")
        rate, ref, cores = cpu_baseline(prob, args.segment)
        oracle_rel = abs(mixed["final_cost"] - ref.final_cost) / abs(ref.final_cost)
        log(f"float64 CPU oracle: {ref.iterations} LM iterations at {rate:.4f} it/s on {cores} cores, final_cost="
            f"{ref.final_cost:.12e}; mixed HIP rel err {oracle_rel:.2e}")
        cpu = {"value": round(rate, 4), "unit": "LM iterations/s", "cores": cores, "kind": "port",
               "sample": f"{ref.iterations} LM iterations (one solve of the timed protocol: Schur-PCG from the same "
                         f"initial point) of the same 500,000-observation workload by the float64 numpy oracle, "
                         f"BLAS allowed {cores} threads (its einsum/scatter work runs on one)"}
    print(json.dumps({
        "metric": "LM iterations/s (mixed-precision HIP Schur-PCG), with matvec HBM GB/s",
        "value": mixed["lm_iters_per_s"], "unit": "LM iterations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * mixed["seconds"] / max(1, mixed["iterations"]), 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 blocks/matvec/PCG vectors, f64 scalars, gradient and objective",
        "data": "synthetic (seeded generator)",
        "config": {"workload": args.workload, "cameras": n_c, "points": n_p, "observations": N, "seed": prob.seed,
                   "linear_solver": "schur_cg", "huber_delta": prob.huber_delta,
                   "protocol": f"{args.steps} LM iterations as solves of {args.segment} from the initial point",
                   "cg_options": CG_OPTIONS},
        "mixed": mixed, "float64": f64,
        "mixed_over_float64": round(mixed["lm_iters_per_s"] / f64["lm_iters_per_s"], 3),
        "mixed_over_float64_per_trial": round(mixed["trials_per_s"] / f64["trials_per_s"], 3),
        "final_cost_rel_err": rel, "final_cost_rel_err_vs_cpu_oracle": oracle_rel,
        "schur_matvec_medium": {"us_per_launch": round(t_schur * 1e6, 2), "bytes_per_launch": schur_bytes,
                                "effective_gbs": round(schur_gbs, 1),
                                "note": "blocks fit the 256 MB last-level cache: effective, not HBM, bandwidth"},
        "first_solve_extra_s": round(plan_s, 4),
        "roofline": roofline, "lm_large": lm5m, "pose_graph": pgraph,
        "dense_schur_cholesky": chol, "implicit_gradient": implicit, "small_cpu_vs_hip": small,
        "cpu_baseline": cpu,
    }), flush=True)


if __name__ == "__main__":
    main()
