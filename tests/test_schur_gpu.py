"""GPU parity of the implicit HIP Schur system and the MFMA Cholesky vs the oracle."""

import math

import numpy as np
import pytest
import torch

from oracle.ba import BAProblem
from oracle.lm import Termination, solve as oracle_solve
from oracle.schur import SchurSystem
from oracle.sparse import NpGroup, NpJacobian
from tests.helpers import rel_err
from torchls_amd import HuberLoss, Problem, SE3Var, TerminationConfig, _lib, generators
from torchls_amd.cholesky import cholesky_f64, cholesky_solve_f64
from torchls_amd.hip_diagonal import LinAlgError
from torchls_amd.residuals import reprojection
from torchls_amd.schur import HipSchurPlan

pytestmark = pytest.mark.gpu


def camera_prior(camera, reference):
    return (camera[:3, :] - reference[:3, :]).reshape(-1)


def point_prior(point, reference):
    return point - reference


def build(prob, device, priors=False):
    P = Problem({"cams": torch.tensor(prob.cameras_init, device=device),
                 "points": torch.tensor(prob.points_init, device=device)}, manifolds={"cams": SE3Var})
    P.add_residual(reprojection, ("cams", torch.tensor(prob.cam_idx)), ("points", torch.tensor(prob.pt_idx)),
                   torch.tensor(prob.observations, device=device), loss=HuberLoss(1e-2))
    if priors:
        P.add_residual(camera_prior, ("cams", 0), torch.tensor(prob.cameras_true[0], device=device))
        P.add_residual(point_prior, ("points", 0), torch.tensor(prob.points_true[0], device=device))
    return P


def to_np(jacobian, layout):
    groups = [NpGroup([b.double().cpu().numpy() for b in g.blocks], list(g.type_indices),
                      [i.cpu().numpy() for i in g.variable_indices]) for g in jacobian.groups]
    return NpJacobian(groups, [(e.count, e.dim, e.offset) for e in layout])


@pytest.mark.parametrize("priors", [False, True])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-11), (torch.float32, 2e-5)])
def test_schur_operators(device, priors, dtype, tol):
    print(f"schur operators priors={priors} dtype={dtype}", flush=True)
    prob = generators.bundle_adjustment(16, 500, 8, seed=21)
    P = build(prob, device, priors)
    an = P._analyze()
    lin = an.linearize(P.parameters)
    diag = tuple(d.double() for d in lin.jacobian.diagonal_blocks())
    lam = 1e-2
    plan = HipSchurPlan(an, "points", device, dtype=dtype)
    system = plan.linearize(lin.jacobian, diag).damped(lam)
    J = to_np(lin.jacobian, an.layout)
    ref = SchurSystem(J, [d.cpu().numpy() for d in diag], 1, lam)
    rng = np.random.default_rng(3)
    rhs = rng.standard_normal(J.n)
    rr = system.reduced_rhs(torch.tensor(rhs, device=device))
    assert rel_err(rr.cpu().numpy(), ref.reduced_rhs(rhs)) <= tol
    x = rng.standard_normal(ref.dim)
    y = system.matvec(torch.tensor(x, device=device, dtype=dtype))
    assert rel_err(y.cpu().numpy(), ref.matvec(x)) <= tol
    full = system.backsubstitute(torch.tensor(x, device=device, dtype=dtype))
    assert rel_err(full.cpu().numpy(), ref.backsubstitute(x)) <= tol
    for a, b in zip(system.preconditioner_blocks(), ref.preconditioner_blocks()):
        assert rel_err(a.cpu().numpy(), b) <= tol
    M = system.dense_matrix().cpu().numpy()
    Md = ref.dense_matrix()
    assert rel_err(np.tril(M), np.tril(Md)) <= tol


@pytest.mark.parametrize("method", ["tiles", "panels"])
@pytest.mark.parametrize("n", [5, 64, 65, 70, 128, 129, 200, 1000, 1536])
def test_cholesky(device, n, method):
    print(f"cholesky n={n} method={method}", flush=True)
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n + 3))
    M = A @ A.T + n * 1e-3 * np.eye(n)
    L = cholesky_f64(torch.tensor(M, device=device), method=method).cpu().numpy()
    Lr = np.linalg.cholesky(M)
    assert rel_err(np.tril(L), Lr) <= 1e-12
    b = rng.standard_normal(n)
    x = cholesky_solve_f64(torch.tensor(M, device=device), torch.tensor(b, device=device), method=method).cpu().numpy()
    assert rel_err(x, np.linalg.solve(M, b)) <= 1e-10
    # the one-workgroup substitution gives the same solution
    xs = torch.tensor(b, device=device)
    Lt = torch.tensor(L, device=device)
    _lib.check(_lib.load().tlh_cholesky_solve_f64(Lt.data_ptr(), n, n, xs.data_ptr(), _lib.stream_handle(device)),
               "tlh_cholesky_solve_f64")
    assert rel_err(x, xs.cpu().numpy()) <= 1e-12
    with pytest.raises(LinAlgError):
        cholesky_f64(torch.tensor(-M, device=device), method=method)
    with pytest.raises(LinAlgError):
        cholesky_solve_f64(torch.tensor(-M, device=device), torch.tensor(b, device=device), method=method)
    # a matrix that fails only at its last pivot
    bad = M.copy()
    bad[n - 1, n - 1] = -1.0
    with pytest.raises(LinAlgError):
        cholesky_f64(torch.tensor(bad, device=device), method=method)


@pytest.mark.parametrize("n", [64, 130, 1536, 4096])
def test_cholesky_tiles_complete(device, n):
    """The persistent tile launch completes without a wait giving up (status 0, so no
    fallback) and writes only the lower triangle; its factor matches LAPACK."""
    print(f"cholesky tiles n={n}", flush=True)
    from torchls_amd.cholesky import _factor

    rng = np.random.default_rng(n + 1)
    A = rng.standard_normal((n, n + 3))
    M = A @ A.T + n * 1e-3 * np.eye(n)
    Mt = torch.tensor(M, device=device)
    L = Mt.clone()
    status = torch.zeros(1, dtype=torch.int32, device=device)
    _factor(L, status, "tiles")
    assert int(status.item()) == 0
    Lh = L.cpu().numpy()
    assert rel_err(np.tril(Lh), np.linalg.cholesky(M)) <= 1e-12
    assert np.array_equal(np.triu(Lh, 1), np.triu(M, 1))
    L2 = Mt.clone()
    _factor(L2, status, "tiles")
    assert torch.equal(L, L2)  # bitwise repeatable


def test_wait_give_up_fallbacks(device):
    """With the bounded waits capped at one poll (tlh_set_wait_limit), the tile
    factorisation gives up (status bit 4) and the blocks solve gives up (bit 2); both
    launches still end, and cholesky_f64 / cholesky_solve_f64 fall back to the panel
    launches and the one-workgroup substitution with results equal to LAPACK's."""
    print("cholesky wait give-up fallbacks n=1000", flush=True)
    from torchls_amd.cholesky import _factor

    lib = _lib.load()
    n = 1000
    rng = np.random.default_rng(7)
    A = rng.standard_normal((n, n + 3))
    M = A @ A.T + n * 1e-3 * np.eye(n)
    b = rng.standard_normal(n)
    Mt, bt = torch.tensor(M, device=device), torch.tensor(b, device=device)
    _lib.check(lib.tlh_set_wait_limit(1), "tlh_set_wait_limit")
    try:
        status = torch.zeros(1, dtype=torch.int32, device=device)
        L = Mt.clone()
        _factor(L, status, "tiles")
        assert int(status.item()) & 4
        status.zero_()
        Lp = Mt.clone()
        _factor(Lp, status, "panels")
        x = bt.clone()
        work = torch.empty(2 * n, dtype=torch.float64, device=device)
        _lib.check(lib.tlh_cholesky_solve_blocks_f64(Lp.data_ptr(), n, n, x.data_ptr(), work.data_ptr(),
                                                     status.data_ptr(), _lib.stream_handle(device)),
                   "tlh_cholesky_solve_blocks_f64")
        assert int(status.item()) == 2
        Lf = cholesky_f64(Mt).cpu().numpy()
        xf = cholesky_solve_f64(Mt, bt).cpu().numpy()
    finally:
        _lib.check(lib.tlh_set_wait_limit(0), "tlh_set_wait_limit")
    assert rel_err(np.tril(Lf), np.linalg.cholesky(M)) <= 1e-12
    assert rel_err(xf, np.linalg.solve(M, b)) <= 1e-10
    assert lib.tlh_set_wait_limit(-1) == _lib.TLH_EINVAL
    # the default limits again: no wait gives up
    status = torch.zeros(1, dtype=torch.int32, device=device)
    L = Mt.clone()
    _factor(L, status, "tiles")
    assert int(status.item()) == 0


def test_cholesky_solve_sentinel_rhs(device):
    """The multi-workgroup solve marks unsolved values with a signalling-NaN bit pattern.
    A right-hand side holding that exact pattern completes end to end (NaN result, no
    wait gives up) and the next solve is unaffected.  (The pattern is quieted by the
    first arithmetic on it, so this checks that the solve completes end to end, not
    the remap branch of publish_value itself.)"""
    print("cholesky sentinel rhs", flush=True)
    n = 200
    rng = np.random.default_rng(7)
    A = rng.standard_normal((n, n + 3))
    M = torch.tensor(A @ A.T + n * 1e-3 * np.eye(n), device=device)
    b = rng.standard_normal(n)
    rhs = torch.tensor(b, device=device)
    rhs.view(torch.int64)[0] = 0x7FF4DEAD5EED1234
    x = cholesky_solve_f64(M, rhs).cpu().numpy()
    assert np.isnan(x).all()
    x = cholesky_solve_f64(M, torch.tensor(b, device=device)).cpu().numpy()
    assert rel_err(x, np.linalg.solve(M.cpu().numpy(), b)) <= 1e-10


@pytest.mark.parametrize("solver", ["schur_cg", "schur_dense"])
def test_schur_lm_matches_oracle(device, solver):
    print(f"schur lm {solver}", flush=True)
    prob = generators.ba_small()
    P = build(prob, device)
    _, s = P.solve(return_summary=True, linear_solver=solver, eliminate="points",
                   termination=TerminationConfig(max_iterations=20))
    o = BAProblem(prob.cam_idx, prob.pt_idx, prob.observations, prob.n_cameras, prob.n_points, huber_delta=1e-2)
    _, ref = oracle_solve(o, (prob.cameras_init, prob.points_init), linear_solver=solver, eliminate=1,
                          term=Termination(max_iterations=20))
    assert s.termination_reason == ref.termination_reason
    assert (s.iterations, s.accepted_steps, s.rejected_steps) == (ref.iterations, ref.accepted_steps, ref.rejected_steps)
    assert math.isclose(s.final_cost, ref.final_cost, rel_tol=1e-9)
    if solver == "schur_cg":
        assert s.linear_iterations_history == ref.linear_iterations_history


def scale_prior(scale, reference):
    return scale - reference


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-11), (torch.float32, 2e-5)])
def test_schur_retained_type_before_cameras(device, dtype, tol):
    """A retained group with only a prior, declared before the cameras, moves the
    camera rows of the reduced system: the dense matrix, matvec and preconditioner
    still match the oracle's closed form."""
    print(f"schur retained type first dtype={dtype}", flush=True)
    prob = generators.bundle_adjustment(16, 500, 8, seed=22)
    rng = np.random.default_rng(5)
    scale0 = rng.standard_normal((3, 2))
    P = Problem({"scale": torch.tensor(scale0, device=device),
                 "cams": torch.tensor(prob.cameras_init, device=device),
                 "points": torch.tensor(prob.points_init, device=device)}, manifolds={"cams": SE3Var})
    P.add_residual(scale_prior, ("scale", torch.arange(3)), torch.tensor(scale0 + 0.3, device=device))
    P.add_residual(reprojection, ("cams", torch.tensor(prob.cam_idx)), ("points", torch.tensor(prob.pt_idx)),
                   torch.tensor(prob.observations, device=device), loss=HuberLoss(1e-2))
    an = P._analyze()
    lin = an.linearize(P.parameters)
    diag = tuple(d.double() for d in lin.jacobian.diagonal_blocks())
    lam = 1e-2
    plan = HipSchurPlan(an, "points", device, dtype=dtype)
    system = plan.linearize(lin.jacobian, diag, dense=True).damped(lam)
    J = to_np(lin.jacobian, an.layout)
    ref = SchurSystem(J, [d.cpu().numpy() for d in diag], 2, lam)
    x = rng.standard_normal(ref.dim)
    y = system.matvec(torch.tensor(x, device=device, dtype=dtype))
    assert rel_err(y.cpu().numpy(), ref.matvec(x)) <= tol
    for a, b in zip(system.preconditioner_blocks(), ref.preconditioner_blocks()):
        assert rel_err(a.cpu().numpy(), b) <= tol
    M = system.dense_matrix().cpu().numpy()
    Md = ref.dense_matrix()
    assert rel_err(np.tril(M), np.tril(Md)) <= tol
    # the LM solve with the dense reduced system matches the Schur-CG solve
    _, s_dense = P.solve(return_summary=True, linear_solver="schur_dense", eliminate="points",
                         termination=TerminationConfig(max_iterations=3))
    assert math.isfinite(s_dense.final_cost) and s_dense.final_cost < s_dense.initial_cost
