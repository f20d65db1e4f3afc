"""GPU parity of the HIP PCG (warp_pcg control flow) against the float64 oracle."""

import math
from functools import partial

import numpy as np
import pytest
import torch

from oracle.pcg import Config, block_jacobi as np_block_jacobi, pcg_batched
from tests.helpers import ba_jacobian, to_block_jacobian
from torchls_amd import _lib, generators
from torchls_amd.hip_diagonal import HipBlockJacobi, HipDiagonalPlan, LinAlgError
from torchls_amd.hip_normal import HipNormalPlan
from torchls_amd.hip_pcg import hip_pcg
from torchls_amd.solvers import ConjugateGradientConfig

pytestmark = pytest.mark.gpu


def _system(device, dtype=torch.float64, seed=5, lam=1e-2, size=(16, 500, 8)):
    prob = generators.bundle_adjustment(*size, seed=seed)
    J, r = ba_jacobian(prob, delta=1e-2)
    BJ = to_block_jacobian(J, device, torch.float64)
    op = HipNormalPlan(BJ, dtype=dtype).bind(BJ)
    diag = HipDiagonalPlan(BJ).diagonal_blocks(BJ)
    g = J.rmatvec(r.reshape(-1))
    return J, BJ, op, diag, g, lam


CONFIGS = [
    dict(),
    dict(residual_check_interval=5, graph_batch_iterations=3),
    dict(restart_strategy="periodic", residual_check_interval=7),
    dict(max_iterations=9),
    dict(residual_check_interval=0, graph_batch_iterations=4),
    dict(rtol=1e-12, max_iterations=60, residual_check_interval=10),
]


@pytest.mark.parametrize("cfg", range(len(CONFIGS)))
def test_matches_oracle_control_flow(device, cfg):
    print(f"pcg control flow cfg={CONFIGS[cfg]}", flush=True)
    J, BJ, op, diag, g, lam = _system(device)
    config = ConjugateGradientConfig(**CONFIGS[cfg])
    pre = HipBlockJacobi(BJ.layout, diag, lam)
    res = hip_pcg(lambda v: op.normal_matvec(v, lam), torch.tensor(-g, device=device), preconditioner=pre,
                  config=config)
    ocfg = Config(**{k: getattr(config, k) for k in Config.__dataclass_fields__})
    ref = pcg_batched(lambda v: J.normal_matvec(v, lam), -g,
                      np_block_jacobi([d.cpu().numpy() for d in diag], J.layout, lam), ocfg)
    assert (res.iterations, res.restarts, res.residual_checks, res.reason, res.converged) == (
        ref.iterations, ref.restarts, ref.residual_checks, ref.reason, ref.converged)
    x = res.solution.cpu().numpy()
    assert np.linalg.norm(x - ref.solution) <= 1e-9 * np.linalg.norm(ref.solution)
    assert math.isclose(res.residual_norm, ref.residual_norm, rel_tol=1e-6, abs_tol=1e-14)


@pytest.mark.parametrize("cfg", [0, 1, 5])
def test_matches_oracle_control_flow_multi_workgroup(device, cfg):
    """n = 18,240 > 16,384: the multi-workgroup dot/step/direction kernels."""
    print(f"pcg control flow (large) cfg={CONFIGS[cfg]}", flush=True)
    J, BJ, op, diag, g, lam = _system(device, size=(40, 6000, 6))
    assert J.n > _lib.PCG_SMALL_MAX_N
    config = ConjugateGradientConfig(**CONFIGS[cfg])
    pre = HipBlockJacobi(BJ.layout, diag, lam)
    res = hip_pcg(partial(op.normal_matvec, damping=lam), torch.tensor(-g, device=device), preconditioner=pre,
                  config=config)
    ocfg = Config(**{k: getattr(config, k) for k in Config.__dataclass_fields__})
    ref = pcg_batched(lambda v: J.normal_matvec(v, lam), -g,
                      np_block_jacobi([d.cpu().numpy() for d in diag], J.layout, lam), ocfg)
    assert (res.iterations, res.restarts, res.residual_checks, res.reason, res.converged) == (
        ref.iterations, ref.restarts, ref.residual_checks, ref.reason, ref.converged)
    x = res.solution.cpu().numpy()
    assert np.linalg.norm(x - ref.solution) <= 1e-9 * np.linalg.norm(ref.solution)


def test_graph_without_warmup_matches_eager(device):
    """A bound HIP operator is captured without the warm-up iteration."""
    J, BJ, op, diag, g, lam = _system(device)
    pre = HipBlockJacobi(BJ.layout, diag, lam)
    rhs = torch.tensor(-g, device=device)
    cfg = ConjugateGradientConfig(residual_check_interval=10, graph_batch_iterations=4)
    mv = partial(op.normal_matvec, damping=lam)
    a = hip_pcg(mv, rhs, preconditioner=pre, config=cfg, use_graph=True)
    b = hip_pcg(mv, rhs, preconditioner=pre, config=cfg, use_graph=False)
    assert torch.equal(a.solution, b.solution) and a.iterations == b.iterations


@pytest.mark.parametrize("size", [(16, 500, 8), (40, 6000, 6)])
@pytest.mark.parametrize("precondition", [True, False])
def test_native_batches_match_graph_and_eager(device, size, precondition):
    """tlh_pcg_iterations issues the captured iteration's launches in the same order:
    native, graph-replayed and eager solves are bitwise equal (small one-workgroup
    and multi-workgroup update paths, fused block Jacobi and none)."""
    print(f"pcg native batches size={size} precondition={precondition}", flush=True)
    from torchls_amd.hip_pcg import native_operator

    J, BJ, op, diag, g, lam = _system(device, size=size)
    pre = HipBlockJacobi(BJ.layout, diag, lam) if precondition else None
    rhs = torch.tensor(-g, device=device)
    cfg = ConjugateGradientConfig(residual_check_interval=10, graph_batch_iterations=3, max_iterations=40)
    mv = partial(op.normal_matvec, damping=lam)
    assert native_operator(mv) is not None and native_operator(lambda v: op.normal_matvec(v, lam)) is None
    a = hip_pcg(mv, rhs, preconditioner=pre, config=cfg, native=True)
    b = hip_pcg(mv, rhs, preconditioner=pre, config=cfg, native=False, use_graph=True)
    c = hip_pcg(mv, rhs, preconditioner=pre, config=cfg, native=False, use_graph=False)
    assert torch.equal(a.solution, b.solution) and torch.equal(a.solution, c.solution)
    assert (a.iterations, a.restarts, a.residual_checks, a.reason) == (
        b.iterations, b.restarts, b.residual_checks, b.reason) == (
        c.iterations, c.restarts, c.residual_checks, c.reason)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_native_schur_batches_match_graph(device, dtype):
    """The Schur system's native batches (W-group pass, other groups' tiles, finish,
    fused update) equal its graph-replayed iterations bitwise, with priors (a second
    retained-only group)."""
    print(f"pcg native schur batches dtype={dtype}", flush=True)
    from tests.test_schur_gpu import build
    from torchls_amd.hip_pcg import native_operator
    from torchls_amd.schur import HipSchurPlan

    prob = generators.bundle_adjustment(16, 500, 8, seed=21)
    P = build(prob, device, priors=True)
    an = P._analyze()
    lin = an.linearize(P.parameters)
    diag = tuple(d.double() for d in lin.jacobian.diagonal_blocks())
    system = HipSchurPlan(an, "points", device, dtype=dtype).linearize(lin.jacobian, diag).damped(1e-2)
    rhs = system.reduced_rhs(lin.jacobian.rmatvec(lin.residual).to(torch.float64), scale=-1.0)
    assert native_operator(system.matvec) is not None
    cfg = ConjugateGradientConfig(residual_check_interval=7, graph_batch_iterations=5)
    a = hip_pcg(system.matvec, rhs, preconditioner=system.preconditioner(), config=cfg, native=True)
    b = hip_pcg(system.matvec, rhs, preconditioner=system.preconditioner(), config=cfg, native=False)
    assert torch.equal(a.solution, b.solution)
    assert (a.iterations, a.restarts, a.residual_checks, a.reason) == (
        b.iterations, b.restarts, b.residual_checks, b.reason)


@pytest.mark.parametrize("dim,count", [(6, 800), (3, 1500)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_single_block_type_step(device, dim, count, dtype):
    """One block-Jacobi type of 3x3 or 6x6 blocks above the one-workgroup size (4,800 /
    4,500 unknowns; the grid-wide dot, step and direction kernels): counts, exit and
    solution match the oracle's batched PCG on the same SPD operator (a block-dominant
    matrix with coupling between blocks), in float64 exactly and in float32 within two
    iterations."""
    print(f"pcg single block type dim={dim} count={count} dtype={dtype}", flush=True)
    from types import SimpleNamespace

    n = dim * count
    rng = np.random.default_rng(dim)
    B = rng.standard_normal((n, 40)) / math.sqrt(40)
    A = B @ B.T
    for k in range(count):  # strong diagonal blocks, different scales
        G = rng.standard_normal((dim, dim))
        A[k * dim:(k + 1) * dim, k * dim:(k + 1) * dim] += (G @ G.T + dim * np.eye(dim)) * (1 + k % 7)
    b = rng.standard_normal(n)
    blocks = A.reshape(count, dim, count, dim)[np.arange(count), :, np.arange(count), :]
    At = torch.tensor(A, device=device, dtype=dtype)
    layout = [SimpleNamespace(count=count, dim=dim, offset=0)]
    pre = HipBlockJacobi(layout, (torch.tensor(blocks, device=device),), 0.0)
    cfg = ConjugateGradientConfig(rtol=1e-8 if dtype == torch.float64 else 1e-5, max_iterations=300,
                                  residual_check_interval=10)
    # eager launches: the operator is a library GEMV, kept out of graph capture
    res = hip_pcg(lambda v: At @ v, torch.tensor(b, device=device, dtype=dtype), preconditioner=pre, config=cfg,
                  use_graph=False)
    ocfg = Config(**{k: getattr(cfg, k) for k in Config.__dataclass_fields__})
    ref = pcg_batched(lambda v: A @ v, b, np_block_jacobi([blocks], [(count, dim, 0)], 0.0), ocfg)
    exact = np.linalg.solve(A, b)
    x = res.solution.double().cpu().numpy()
    if dtype == torch.float64:
        assert (res.iterations, res.restarts, res.residual_checks, res.reason) == (
            ref.iterations, ref.restarts, ref.residual_checks, ref.reason)
        assert np.linalg.norm(x - ref.solution) <= 1e-9 * np.linalg.norm(ref.solution)
    else:
        assert res.reason == "residual_tolerance" and abs(res.iterations - ref.iterations) <= 2
    assert np.linalg.norm(x - exact) <= (1e-6 if dtype == torch.float64 else 1e-3) * np.linalg.norm(exact)


def test_deferred_positive_definiteness_check(device):
    J, BJ, op, diag, g, lam = _system(device)
    bad = tuple(-d.double() for d in diag)
    pre = HipBlockJacobi(BJ.layout, bad, 0.0, defer_check=True)  # no host sync, no raise yet
    with pytest.raises(LinAlgError):
        hip_pcg(partial(op.normal_matvec, damping=lam), torch.tensor(-g, device=device), preconditioner=pre,
                config=ConjugateGradientConfig())
    with pytest.raises(LinAlgError):
        HipBlockJacobi(BJ.layout, bad, 0.0)


def test_graph_replay_matches_eager(device):
    J, BJ, op, diag, g, lam = _system(device)
    pre = HipBlockJacobi(BJ.layout, diag, lam)
    rhs = torch.tensor(-g, device=device)
    cfg = ConjugateGradientConfig(residual_check_interval=10, graph_batch_iterations=4)
    a = hip_pcg(lambda v: op.normal_matvec(v, lam), rhs, preconditioner=pre, config=cfg, use_graph=True)
    b = hip_pcg(lambda v: op.normal_matvec(v, lam), rhs, preconditioner=pre, config=cfg, use_graph=False)
    c = hip_pcg(lambda v: op.normal_matvec(v, lam), rhs, preconditioner=pre, config=cfg, use_graph=True)
    assert torch.equal(a.solution, b.solution) and torch.equal(a.solution, c.solution)
    assert a.iterations == b.iterations == c.iterations


@pytest.mark.parametrize("kind", ["none", "callable"])
def test_other_preconditioners(device, kind):
    J, BJ, op, diag, g, lam = _system(device)
    rhs = torch.tensor(-g, device=device)
    if kind == "none":
        pre, npre = None, None
    else:
        dvec = torch.cat([d.diagonal(dim1=-2, dim2=-1).reshape(-1) for d in diag]) + lam
        pre = lambda v: v / dvec  # noqa: E731
        dn = dvec.cpu().numpy()
        npre = lambda v: v / dn  # noqa: E731
    cfg = ConjugateGradientConfig(max_iterations=300)
    res = hip_pcg(lambda v: op.normal_matvec(v, lam), rhs, preconditioner=pre, config=cfg)
    ocfg = Config(max_iterations=300)
    ref = pcg_batched(lambda v: J.normal_matvec(v, lam), -g, npre, ocfg)
    assert (res.iterations, res.restarts, res.reason) == (ref.iterations, ref.restarts, ref.reason)
    assert np.linalg.norm(res.solution.cpu().numpy() - ref.solution) <= 1e-8 * np.linalg.norm(ref.solution)


def test_float32_mixed(device):
    J, BJ, op32, diag, g, lam = _system(device, dtype=torch.float32)
    pre = HipBlockJacobi(BJ.layout, diag, lam)
    res = hip_pcg(lambda v: op32.normal_matvec(v, lam), torch.tensor(-g, device=device, dtype=torch.float32),
                  preconditioner=pre, config=ConjugateGradientConfig(rtol=1e-5))
    A = J.to_dense()
    exact = np.linalg.solve(A.T @ A + lam * np.eye(J.n), -g)
    assert res.reason == "residual_tolerance"
    assert np.linalg.norm(res.solution.cpu().numpy() - exact) <= 1e-3 * np.linalg.norm(exact)


def test_failure_reasons(device):
    J, BJ, op, diag, g, lam = _system(device)
    rhs = torch.tensor(-g, device=device)
    cfg = ConjugateGradientConfig()
    nan = rhs.clone()
    nan[3] = float("nan")
    r = hip_pcg(lambda v: op.normal_matvec(v, lam), nan, preconditioner=None, config=cfg)
    assert r.reason == "nonfinite" and not r.converged and r.iterations == 0
    z = hip_pcg(lambda v: op.normal_matvec(v, lam), torch.zeros_like(rhs), preconditioner=None, config=cfg)
    assert z.reason == "residual_tolerance" and z.converged and z.iterations == 0
    neg = hip_pcg(lambda v: -op.normal_matvec(v, lam), rhs, preconditioner=None, config=cfg)
    ref = pcg_batched(lambda v: -J.normal_matvec(v, lam), -g, None, Config())
    assert neg.reason == ref.reason == "nonpositive_curvature"
    assert neg.iterations == ref.iterations


@pytest.mark.parametrize("case", range(11))
def test_counts_by_hand(device, case):
    """The hand-derived warp_pcg counts of the oracle pins, on the HIP PCG."""
    from tests.test_oracle_pins import PCG_COUNT_CASES, pcg_count_system

    k, cfg, expected = PCG_COUNT_CASES[case]
    print(f"pcg counts by hand k={k} cfg={cfg}", flush=True)
    eig, b = pcg_count_system(k)
    e = torch.tensor(eig, device=device)
    res = hip_pcg(lambda v: e * v, torch.tensor(b, device=device), preconditioner=None,
                  config=ConjugateGradientConfig(**cfg))
    assert (res.iterations, res.restarts, res.residual_checks, res.reason, res.converged) == expected
    if res.converged:
        assert np.allclose(res.solution.cpu().numpy(), b / eig, rtol=1e-12)


def test_exits_by_hand(device):
    from tests.test_oracle_pins import pcg_count_system

    eig, b = pcg_count_system(5)
    e = torch.tensor(eig, device=device)
    bt = torch.tensor(b, device=device)
    cfg = ConjugateGradientConfig()
    res = hip_pcg(lambda v: e * v, bt, preconditioner=lambda v: v / e, config=cfg)
    assert (res.iterations, res.restarts, res.residual_checks, res.reason, res.converged) == (
        1, 0, 1, "residual_tolerance", True)
    d = torch.tensor([1.0, -1.0], device=device)
    res = hip_pcg(lambda v: d * v, torch.tensor([1.0, 2.0], device=device), preconditioner=None, config=cfg)
    assert (res.iterations, res.restarts, res.residual_checks, res.reason, res.converged) == (
        1, 0, 1, "nonpositive_curvature", False)
    assert torch.count_nonzero(res.solution).item() == 0
    res = hip_pcg(lambda v: v * float("nan"), bt, preconditioner=None, config=cfg)
    assert (res.iterations, res.restarts, res.residual_checks, res.reason, res.converged) == (
        1, 0, 1, "nonfinite", False)


@pytest.mark.parametrize("limit", [0, 2])
def test_refinement_limit_matches_oracle(device, limit):
    """max_refinements on a float32-rounded diagonal operator (elementwise products
    round identically here and in numpy): the oracle's counts and exit."""
    from tests.test_oracle_pins import refinement_system

    _, e, b = refinement_system()
    e32 = torch.tensor(e, device=device).float()
    cfg = dict(rtol=1e-12, max_iterations=1000, residual_check_interval=0, residual_drift_tolerance=1e30,
               max_refinements=limit)
    res = hip_pcg(lambda v: (e32 * v.float()).double(), torch.tensor(b, device=device), preconditioner=None,
                  config=ConjugateGradientConfig(**cfg))
    matvec, _, _ = refinement_system()
    ref = pcg_batched(matvec, b, None, Config(**cfg))
    assert (res.restarts, res.residual_checks, res.reason) == (limit, limit + 1, "stagnated")
    assert (res.iterations, res.restarts, res.residual_checks, res.reason) == (
        ref.iterations, ref.restarts, ref.residual_checks, ref.reason)



def test_drift_restarts_off_matches_oracle(device):
    """drift_restarts=False on the float32-rounded diagonal operator, checks every 10
    iterations: the oracle's counts and exit (restarts only at converged recurrences)."""
    from tests.test_oracle_pins import refinement_system

    _, e, b = refinement_system()
    e32 = torch.tensor(e, device=device).float()
    cfg = dict(rtol=1e-12, max_iterations=200, residual_check_interval=10, drift_restarts=False)
    res = hip_pcg(lambda v: (e32 * v.float()).double(), torch.tensor(b, device=device), preconditioner=None,
                  config=ConjugateGradientConfig(**cfg))
    matvec, _, _ = refinement_system()
    ref = pcg_batched(matvec, b, None, Config(**cfg))
    assert (res.iterations, res.restarts, res.residual_checks, res.reason) == (
        ref.iterations, ref.restarts, ref.residual_checks, ref.reason)
