"""Benchmark of the HIP block-sparse normal-equation path on one MI355X.

Prints one JSON line (see --help).  The timed region holds only device work on
inputs already resident in HBM; a bounded CPU-oracle sample is timed beside it.
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


def parse():
    p = argparse.ArgumentParser(description=__doc__)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--workload", default="ba_large", choices=["ba_medium", "ba_large"])
    p.add_argument("--no-cpu", action="store_true")
    return p.parse_args()


def cpu_baseline_matvec(prob, seconds=10.0):
    """Oracle (numpy float64) normal matvec on the same workload, host cores."""
    from oracle.sparse import NpGroup, NpJacobian
    from oracle.ba import linearize_ba

    r, Jc, Jp, _, _ = linearize_ba(prob.cameras_init, prob.points_init, prob.cam_idx, prob.pt_idx,
                                   prob.observations, prob.huber_delta)
    J = NpJacobian([NpGroup([Jc, Jp], [0, 1], [prob.cam_idx, prob.pt_idx])],
                   [(prob.n_cameras, 6, 0), (prob.n_points, 3, 6 * prob.n_cameras)])
    x = np.random.default_rng(0).standard_normal(J.n)
    J.normal_matvec(x, 1e-3)
    reps, t0 = 0, time.perf_counter()
    while True:
        J.normal_matvec(x, 1e-3)
        reps += 1
        if time.perf_counter() - t0 >= seconds or reps >= 50:
            break
    return (time.perf_counter() - t0) / reps


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    from torchls_amd import generators
    from torchls_amd.hip_normal import HipNormalPlan
    from torchls_amd.workloads import ba_linearization, normal_matvec_bytes

    device = torch.device("cuda:0")
    prob = getattr(generators, args.workload)()
    print(f"workload {args.workload}: cams={prob.n_cameras} points={prob.n_points} obs={prob.n_obs}",
          file=sys.stderr, flush=True)
    res, J = ba_linearization(prob, device, torch.float64)
    plan = HipNormalPlan(J, dtype=torch.float32)
    op = plan.bind(J)
    n = J.shape[1]
    x = torch.randn(n, dtype=torch.float32, device=device)
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        op.normal_matvec_into(x, y, 1e-3)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(args.steps):
        op.normal_matvec_into(x, y, 1e-3)
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / args.steps
    bytes_mv = normal_matvec_bytes(prob.n_obs, 2, (6, 3), n, 4)
    gbs = bytes_mv / (ms * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        t = cpu_baseline_matvec(prob)
        cpu = {"value": round(normal_matvec_bytes(prob.n_obs, 2, (6, 3), n, 8) / t / 1e9, 3),
               "unit": "GB/s (float64 normal matvec)", "cores": 1, "kind": "port",
               "sample": f"{args.workload} full size, numpy float64 oracle normal matvec"}
    print(json.dumps({
        "metric": "normal matvec HBM GB/s (float32 blocks, deterministic reduction)",
        "value": round(gbs, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": args.workload, "cameras": prob.n_cameras, "points": prob.n_points,
                   "observations": prob.n_obs},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": cpu,
    }), flush=True)


if __name__ == "__main__":
    main()
