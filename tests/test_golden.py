"""The oracle (and the host-side loss) against golden vectors of the reference torch path.

The vectors in tests/golden/*.npz were written by tests/golden/make_golden.py from
torchls's own code (solvers.pcg, solvers.solve, HuberLoss) on CPU float64.
"""

import os

import numpy as np
import pytest
import torch

from oracle.ba import BAProblem, huber_reduction, huber_value
from oracle.lm import TrustRegion, Termination, solve as oracle_solve
from oracle.pcg import Config, pcg_reference
from torchls_amd.losses import HuberLoss

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def test_huber_oracle_and_host_loss():
    g = load("huber.npz")
    d = float(g["delta"])
    assert np.allclose(huber_value(g["a"], d), g["value"], rtol=1e-15, atol=0)
    assert np.allclose(huber_reduction(g["a"], g["b"], g["diff"], d), g["reduction"], rtol=1e-13, atol=1e-300)
    loss = HuberLoss(d)
    v, f, s = loss.evaluate(torch.tensor(g["a"]))
    assert np.allclose(v.numpy(), g["value"], rtol=1e-15)
    assert np.allclose(f.numpy(), g["first"], rtol=1e-15)
    assert np.allclose(s.numpy(), g["second"], rtol=1e-15)
    r = loss._reduction(torch.tensor(g["a"]), torch.tensor(g["b"]), torch.tensor(g["diff"]))
    assert np.allclose(r.numpy(), g["reduction"], rtol=1e-13, atol=1e-300)


@pytest.mark.parametrize("case", range(6))
@pytest.mark.parametrize("pre", [0, 1])
def test_pcg_reference_counts(case, pre):
    g = load("pcg.npz")
    cfg = eval(str(g["configs"][case]), {})  # dict literal written by make_golden.py
    M, b, d = g["M"], g["b"], g["d"]
    r = pcg_reference(lambda v: M @ v, b, (lambda v: v / d) if pre else None, Config(**cfg))
    key = f"c{case}_p{pre}"
    assert [r.iterations, r.restarts, r.residual_checks, int(r.converged)] == list(g[key + "_counts"])
    assert r.reason == str(g[key + "_reason"])
    # iterates agree to the solve's own accuracy (rounding differs between torch and numpy dots)
    assert np.linalg.norm(r.solution - g[key + "_x"]) <= 1e-6 * np.linalg.norm(g[key + "_x"])


CASES = {
    "cg": dict(),
    "cg_adaptive": dict(tr=TrustRegion(strategy="adaptive")),
    "schur_cg": dict(linear_solver="schur_cg", eliminate=1),
    "schur_dense": dict(linear_solver="schur_dense", eliminate=1),
}


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_lm_trajectory(name):
    g = load("lm.npz")
    prob = BAProblem(g["cam_idx"], g["pt_idx"], g["uv"], g["cams0"].shape[0], g["pts0"].shape[0], huber_delta=3e-3)
    (cams, pts), s = oracle_solve(prob, (g["cams0"], g["pts0"]), term=Termination(max_iterations=20),
                                  pcg_flavour="reference", **CASES[name])
    assert [s.iterations, s.accepted_steps, s.rejected_steps] == list(g[name + "_counts"])
    assert s.termination_reason == str(g[name + "_reason"])
    assert np.allclose(s.cost_history, g[name + "_cost"], rtol=1e-9, atol=0)
    assert np.allclose(s.lambda_history, g[name + "_lambda"], rtol=1e-12, atol=0)
    if not name.endswith("dense"):
        assert list(s.linear_iterations_history) == list(g[name + "_linear"])
    assert np.allclose(pts, g[name + "_pts"], rtol=0, atol=1e-6)
