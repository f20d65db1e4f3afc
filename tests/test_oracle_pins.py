"""Pins of the float64 oracle itself, by closed forms, brute force and textbook routines.

Each check uses code written here, independent of the oracle function it pins, so
that a plausible mistake in the oracle (a dropped alias term, a transposed block,
a wrong sign or a missing damping term) fails at least one assertion.
"""

from fractions import Fraction

import math

import numpy as np
import pytest
import scipy.linalg

from oracle.ba import huber_reduction, huber_value, huber_weights, linearize_ba, reprojection
from oracle.pcg import Config, block_jacobi, pcg_batched, pcg_reference
from oracle.schur import SchurSystem
from oracle.sparse import NpGroup, NpJacobian, random_jacobian
from torchls_amd import generators

LAYOUT = [(6, 6), (40, 3), (3, 1)]
GROUPS = [(120, 2, [0, 1]), (15, 3, [1]), (30, 6, [0, 0]), (10, 2, [2, 1, 2])]


def brute_dense(J):
    """Dense J by explicit loops over observations, slots, rows and columns."""
    A = np.zeros((J.m, J.n))
    row0 = 0
    for g in J.groups:
        for o in range(g.count):
            for s, t in enumerate(g.types):
                count, dim, off = J.layout[t]
                for r in range(g.rows):
                    for c in range(dim):
                        A[row0 + o * g.rows + r, off + g.indices[s][o] * dim + c] += g.blocks[s][o, r, c]
        row0 += g.count * g.rows
    return A


@pytest.fixture(scope="module")
def jac():
    return random_jacobian(np.random.default_rng(11), LAYOUT, GROUPS)


def test_sparse_operators_against_brute_force(jac):
    A = brute_dense(jac)
    rng = np.random.default_rng(2)
    x, y = rng.standard_normal(jac.n), rng.standard_normal(jac.m)
    assert np.allclose(jac.to_dense(), A, rtol=0, atol=0)
    assert np.allclose(jac.matvec(x), A @ x, rtol=1e-13, atol=1e-13)
    assert np.allclose(jac.rmatvec(y), A.T @ y, rtol=1e-13, atol=1e-13)
    assert np.allclose(jac.normal_matvec(x, 0.3), A.T @ (A @ x) + 0.3 * x, rtol=1e-12, atol=1e-12)
    assert np.isclose(jac.squared_matvec_norm(x), np.sum((A @ x) ** 2), rtol=1e-12)
    s = rng.uniform(0.5, 2.0, jac.n)
    assert np.allclose(jac.scale_columns(s).to_dense(), A * s[None, :], rtol=1e-14, atol=0)


def test_diagonal_blocks_include_alias_terms(jac):
    A = brute_dense(jac)
    N = A.T @ A
    blocks = jac.diagonal_blocks()
    for (count, dim, off), D in zip(jac.layout, blocks):
        for v in range(count):
            sl = slice(off + v * dim, off + (v + 1) * dim)
            assert np.allclose(D[v], N[sl, sl], rtol=1e-12, atol=1e-12)


def _se3_twist(xi):
    v, w = xi[:3], xi[3:]
    m = np.zeros((4, 4))
    m[:3, :3] = [[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]]
    m[:3, 3] = v
    return m


def test_reprojection_blocks_are_the_derivatives():
    """Central differences of the residual through T expm(hat(delta)) and p + delta."""
    prob = generators.bundle_adjustment(4, 6, 3, seed=9)
    cams, pts = prob.cameras_init, prob.points_init
    ci, pi, uv = prob.cam_idx, prob.pt_idx, prob.observations
    _, Jc, Jp = reprojection(cams, pts, ci, pi, uv)
    h = 1e-6
    for o in range(len(ci)):
        T, p = cams[ci[o]], pts[pi[o]]

        def res(Tm, pm):
            q = Tm[:3, :3] @ pm + Tm[:3, 3]
            return q[:2] / q[2] - uv[o]

        for k in range(6):
            e = np.zeros(6)
            e[k] = h
            fd = (res(T @ scipy.linalg.expm(_se3_twist(e)), p) - res(T @ scipy.linalg.expm(_se3_twist(-e)), p)) / (2 * h)
            assert np.allclose(Jc[o][:, k], fd, rtol=1e-6, atol=1e-8)
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            fd = (res(T, p + e) - res(T, p - e)) / (2 * h)
            assert np.allclose(Jp[o][:, k], fd, rtol=1e-6, atol=1e-8)


def test_huber_weighting_closed_form():
    prob = generators.bundle_adjustment(4, 30, 3, seed=3)
    delta = 3e-3
    model, Jc, Jp, raw, half = linearize_ba(prob.cameras_init, prob.points_init, prob.cam_idx, prob.pt_idx,
                                            prob.observations, delta)
    s = np.sum(raw ** 2, -1)
    w = np.where(np.sqrt(s) > delta, np.sqrt(delta / np.sqrt(s)), 1.0)
    assert np.allclose(model, raw * w[:, None], rtol=1e-15)
    rho = np.where(np.sqrt(s) > delta, 2 * delta * np.sqrt(s) - delta ** 2, s)
    assert np.isclose(half, 0.5 * rho.sum(), rtol=1e-14)
    sw, v = huber_weights(raw, delta)
    assert np.allclose(sw, w, rtol=1e-15) and np.allclose(v, rho, rtol=1e-15)


@pytest.mark.parametrize("a,b", [(0.5625, 0.25), (4.0, 9.0), (0.25, 16.0), (16.0, 0.25), (2.25, 2.25)])
def test_huber_reduction_is_the_exact_difference(a, b):
    """rho(a) - rho(b) in exact rational arithmetic (delta = 1, perfect squares)."""
    def rho(s):
        r = Fraction(s).limit_denominator()
        root = Fraction(int(round(float(r) ** 0.5 * 4)), 4)
        assert root * root == r
        return r if root <= 1 else 2 * root - 1

    exact = float(rho(a) - rho(b))
    diff = a - b  # one-row residuals: (x - y)(x + y) = a - b
    got = float(huber_reduction(np.array([a]), np.array([b]), np.array([diff]), 1.0)[0])
    assert abs(got - exact) <= 1e-15 * max(1.0, abs(exact))
    assert float(huber_value(np.array([a]), 1.0)[0]) == float(rho(a))


def _ba_system(damping):
    prob = generators.bundle_adjustment(5, 25, 4, seed=21)
    _, Jc, Jp, _, _ = linearize_ba(prob.cameras_init, prob.points_init, prob.cam_idx, prob.pt_idx,
                                   prob.observations, None)
    layout = [(5, 6, 0), (25, 3, 30)]
    return NpJacobian([NpGroup([Jc, Jp], [0, 1], [prob.cam_idx, prob.pt_idx])], layout)


def test_schur_against_dense_closed_form():
    lam = 0.02
    J = _ba_system(lam)
    A = brute_dense(J)
    H = A.T @ A + lam * np.eye(J.n)
    c, e = slice(0, 30), slice(30, J.n)
    S = H[c, c] - H[c, e] @ np.linalg.solve(H[e, e], H[e, c])
    sysm = SchurSystem(J, J.diagonal_blocks(), 1, lam)
    rng = np.random.default_rng(4)
    x = rng.standard_normal(30)
    assert np.allclose(sysm.matvec(x), S @ x, rtol=1e-10, atol=1e-10)
    assert np.allclose(sysm.dense_matrix(), S, rtol=1e-10, atol=1e-10)
    for v, M in enumerate(sysm.preconditioner_blocks()[0]):
        assert np.allclose(M, S[6 * v: 6 * v + 6, 6 * v: 6 * v + 6], rtol=1e-10, atol=1e-10)
    rhs = rng.standard_normal(J.n)
    xr = np.linalg.solve(S, sysm.reduced_rhs(rhs))
    assert np.allclose(sysm.backsubstitute(xr), np.linalg.solve(H, rhs), rtol=1e-9, atol=1e-10)


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((n, n))
    return B @ B.T + n * np.eye(n), rng.standard_normal(n)


def test_pcg_reference_is_textbook_pcg():
    M, b = _spd(30, 1)
    d = np.diag(M).copy()
    res = pcg_reference(lambda v: M @ v, b, lambda v: v / d, Config(rtol=1e-12, atol=0, max_iterations=200))
    assert res.converged and np.allclose(res.solution, np.linalg.solve(M, b), rtol=1e-9)
    # first three iterates of the Hestenes-Stiefel recurrence (Saad, Alg. 9.1)
    x = np.zeros(30)
    r = b.copy()
    z = r / d
    p = z.copy()
    for _ in range(3):
        a = (r @ z) / (p @ (M @ p))
        x = x + a * p
        r_new = r - a * (M @ p)
        z_new = r_new / d
        p = z_new + (r_new @ z_new) / (r @ z) * p
        r, z = r_new, z_new
    three = pcg_reference(lambda v: M @ v, b, lambda v: v / d, Config(rtol=0, atol=0, max_iterations=3))
    assert np.allclose(three.solution, x, rtol=1e-12, atol=1e-14) and three.iterations == 3


def test_pcg_batched_matches_reference_without_checks():
    M, b = _spd(40, 2)
    cfg = Config(rtol=0, atol=0, max_iterations=12, residual_check_interval=0)
    a = pcg_reference(lambda v: M @ v, b, None, cfg)
    c = pcg_batched(lambda v: M @ v, b, None, cfg)
    assert a.iterations == c.iterations == 12
    assert np.allclose(a.solution, c.solution, rtol=1e-12, atol=1e-14)
    full = pcg_batched(lambda v: M @ v, b, None, Config(rtol=1e-12, atol=0, max_iterations=400))
    assert full.converged and np.allclose(full.solution, np.linalg.solve(M, b), rtol=1e-9)


def test_block_jacobi_is_the_blockwise_solve(jac):
    blocks = jac.diagonal_blocks()
    apply = block_jacobi(blocks, jac.layout, 0.7)
    v = np.random.default_rng(8).standard_normal(jac.n)
    out = apply(v)
    for (count, dim, off), D in zip(jac.layout, blocks):
        for k in range(count):
            sl = slice(off + k * dim, off + (k + 1) * dim)
            assert np.allclose(out[sl], np.linalg.solve(D[k] + 0.7 * np.eye(dim), v[sl]), rtol=1e-10, atol=1e-12)


# ---------------------------------------------------------------- warp_pcg counts by hand
# Diagonal SPD systems whose eigenvalues take k distinct powers of two, with a
# right-hand side of ones: unpreconditioned CG reaches the exact solution after k
# iterations (to rounding, far below the 1e-6 relative threshold) and not before.
# The (iterations, restarts, checks, reason, converged) below are derived by hand
# from the batched control flow of warp_pcg.py:176-213: a batch replays up to the
# window end stop_at = min(max_iterations, next multiple of residual_check_interval);
# the host checks the true residual when the status is nonzero or the window ended;
# a check restarts on status 1 / periodic strategy / drift (never here: the drift is
# rounding-sized) unless the residual converged or max_iterations was reached.
PCG_COUNT_CASES = [
    # (k distinct eigenvalues, config, expected)
    (3, dict(), (3, 0, 1, "residual_tolerance", True)),            # one check, at status 1
    (5, dict(residual_check_interval=2), (5, 0, 3, "residual_tolerance", True)),  # windows end at 2, 4; status 1 at 5
    (5, dict(residual_check_interval=2, graph_batch_iterations=1), (5, 0, 3, "residual_tolerance", True)),
    (5, dict(residual_check_interval=2, graph_batch_iterations=3), (5, 0, 3, "residual_tolerance", True)),
    (5, dict(residual_check_interval=0, graph_batch_iterations=4), (5, 0, 1, "residual_tolerance", True)),
    (4, dict(residual_check_interval=4), (4, 0, 1, "residual_tolerance", True)),  # status 1 at the window end
    (5, dict(residual_check_interval=2, max_iterations=3), (3, 0, 2, "max_iterations", False)),
    (5, dict(residual_check_interval=2, max_iterations=3, restart_strategy="periodic"),
     (3, 1, 2, "max_iterations", False)),                           # restart at the check of iteration 2
    (5, dict(max_iterations=5), (5, 0, 1, "residual_tolerance", True)),
    (5, dict(max_iterations=4), (4, 0, 1, "max_iterations", False)),
    (5, dict(residual_check_interval=3, restart_strategy="periodic", max_iterations=7),
     (7, 2, 3, "max_iterations", False)),                           # restarts at 3 and 6, last check at 7
]


def pcg_count_system(k, n=20):
    eig = 2.0 ** (np.arange(n) % k)
    return eig, np.ones(n)


@pytest.mark.parametrize("case", range(len(PCG_COUNT_CASES)))
def test_pcg_batched_counts_by_hand(case):
    k, cfg, expected = PCG_COUNT_CASES[case]
    eig, b = pcg_count_system(k)
    res = pcg_batched(lambda v: eig * v, b, None, Config(**cfg))
    assert (res.iterations, res.restarts, res.residual_checks, res.reason, res.converged) == expected
    if res.converged:
        assert np.allclose(res.solution, b / eig, rtol=1e-12)


def test_pcg_batched_exits_by_hand():
    """Exits at the first read: an exact block-Jacobi preconditioner converges in one
    iteration; an indefinite diagonal stops at status 3 after one (frozen) iteration;
    a NaN operator at status 2; a zero or NaN right-hand side before any iteration."""
    eig, b = pcg_count_system(5)
    res = pcg_batched(lambda v: eig * v, b, lambda v: v / eig, Config())
    assert (res.iterations, res.restarts, res.residual_checks, res.reason, res.converged) == (
        1, 0, 1, "residual_tolerance", True)
    res = pcg_batched(lambda v: np.array([1.0, -1.0]) * v, np.array([1.0, 2.0]), None, Config())
    assert (res.iterations, res.restarts, res.residual_checks, res.reason, res.converged) == (
        1, 0, 1, "nonpositive_curvature", False)
    assert np.array_equal(res.solution, np.zeros(2))  # the guarded step was not taken
    res = pcg_batched(lambda v: v * np.nan, b, None, Config())
    assert (res.iterations, res.restarts, res.residual_checks, res.reason, res.converged) == (
        1, 0, 1, "nonfinite", False)
    res = pcg_batched(lambda v: eig * v, np.zeros_like(b), None, Config())
    assert (res.iterations, res.residual_checks, res.reason, res.converged) == (0, 0, "residual_tolerance", True)
    res = pcg_batched(lambda v: eig * v, b * np.nan, None, Config())
    assert (res.iterations, res.residual_checks, res.reason, res.converged) == (0, 0, "nonfinite", False)


def refinement_system():
    """Diagonal operator rounded to float32 (e * v with both factors rounded): its
    true residual cannot fall below ~1e-7 |b|, far above rtol 1e-12, while the
    float64 recurrence keeps converging."""
    rng = np.random.default_rng(0)
    e = np.linspace(1.0, 100.0, 40)
    e32 = e.astype(np.float32)
    return (lambda v: (e32 * v.astype(np.float32)).astype(np.float64)), e, rng.standard_normal(40)


@pytest.mark.parametrize("limit", [0, 1, 2])
def test_pcg_batched_refinement_limit(limit):
    """max_refinements=k: with drift restarts disabled and checks only at a converged
    recurrence, exactly k restarts (k + 1 checks) and the exit "stagnated"; the same
    operator without rounding converges at the first check."""
    matvec, e, b = refinement_system()
    cfg = dict(rtol=1e-12, max_iterations=1000, residual_check_interval=0, residual_drift_tolerance=1e30)
    res = pcg_batched(matvec, b, None, Config(max_refinements=limit, **cfg))
    assert (res.restarts, res.residual_checks, res.reason, res.converged) == (limit, limit + 1, "stagnated", False)
    assert 1e-9 < res.residual_norm / np.linalg.norm(b) < 1e-6
    exact = pcg_batched(lambda v: e * v, b, None, Config(max_refinements=0, **cfg))
    assert (exact.restarts, exact.residual_checks, exact.reason) == (0, 1, "residual_tolerance")
    # without a limit (and without stop_on_stagnation) the solve restarts until max_iterations
    free = pcg_batched(matvec, b, None, Config(**dict(cfg, max_iterations=300)))
    assert free.reason == "max_iterations" and free.restarts > limit


def test_pcg_batched_drift_restarts_off():
    """drift_restarts=False removes exactly the drift clause of the restart test: on
    the float32-rounded operator (its true residual floors near 1e-7 |b|, so checks
    see drift) the solve equals the one with an unbounded drift tolerance, bit for
    bit, and restarts only at converged recurrences; with drift restarts it restarts
    at more checks."""
    matvec, _, b = refinement_system()
    cfg = dict(rtol=1e-12, max_iterations=200, residual_check_interval=10)
    off = pcg_batched(matvec, b, None, Config(drift_restarts=False, **cfg))
    unbounded = pcg_batched(matvec, b, None, Config(residual_drift_tolerance=1e30, **cfg))
    assert (off.iterations, off.restarts, off.residual_checks, off.reason) == (
        unbounded.iterations, unbounded.restarts, unbounded.residual_checks, unbounded.reason)
    assert np.array_equal(off.solution, unbounded.solution)
    drifting = pcg_batched(matvec, b, None, Config(**cfg))
    assert drifting.restarts > off.restarts


@pytest.mark.parametrize("mode", ["diagonal", "jacobi", "both"])
def test_lm_scaling_is_marquardt_damping(mode):
    """Column scaling s and lambda I on J diag(s) is the damped system
    (J^T J + lambda diag(s)^-2) delta = -g of the unscaled problem: Marquardt's
    lambda diag(J^T J) for damping_mode="diagonal", lambda (1 + sqrt(diag))^2 for
    Jacobi scaling.  The first accepted step of the oracle LM equals that solve."""
    from oracle.ba import BAProblem
    from oracle.lm import Termination, TrustRegion, solve as lm_solve
    from torchls_amd.generators import bundle_adjustment

    prob = bundle_adjustment(6, 40, 4, seed=31)
    o = BAProblem(prob.cam_idx, prob.pt_idx, prob.observations, prob.n_cameras, prob.n_points, huber_delta=None)
    vals0 = (prob.cameras_init, prob.points_init)
    tr = TrustRegion(damping_mode="diagonal" if mode in ("diagonal", "both") else "identity")
    vals1, s = lm_solve(o, vals0, linear_solver="dense", tr=tr, term=Termination(max_iterations=1),
                        jacobi_scaling=mode in ("jacobi", "both"))
    assert (s.accepted_steps, s.rejected_steps) == (1, 0)
    lin = o.linearize(vals0)
    A = lin["J"].to_dense()
    H = A.T @ A
    g = A.T @ lin["residual"]
    d = np.diag(H).copy()
    if mode == "diagonal":
        damp = np.clip(d, tr.diagonal_min, tr.diagonal_max)
    elif mode == "jacobi":
        damp = (1.0 + np.sqrt(d)) ** 2
    else:
        s2 = 1.0 / (1.0 + np.sqrt(d)) ** 2
        damp = np.clip(d * s2, tr.diagonal_min, tr.diagonal_max) / s2
    delta = np.linalg.solve(H + tr.lambda_initial * np.diag(damp), -g)
    expect = o.retract(vals0, delta)
    for a, b in zip(vals1, expect):
        assert np.allclose(a, b, rtol=1e-9, atol=1e-12)


def test_trial_decision_by_hand():
    """oracle.lm.trial_decision (solvers.py:607-617, 633-655, 683-693) on cases whose
    outcome follows from the rules by hand: the factor rule divides lambda by the
    factor above a gain ratio of 0.75 and multiplies it on a rejection; the adaptive
    rule scales lambda by max(1/3, 1 - (2 min(q, 1) - 1)^3) (1/3 at q = 1, 1 at q = 1/2,
    7/8 at q = 3/4) and multiplies it by a rejection factor that doubles up to
    lambda_max / lambda_min; a nonpositive predicted or actual reduction, a gain ratio
    below the threshold or a nonfinite cost rejects; a rejection at lambda_max, or
    without a trust region, stops the trials."""
    from oracle.lm import TrustRegion, trial_decision

    fixed = TrustRegion(lambda_min=1e-6, lambda_max=1.0, lambda_factor=4.0, step_quality_min=1e-4)
    assert trial_decision(1.0, 0.9, 1.0, 0.5, 2.0, fixed) == (0.9, True, 0.125, 2.0, False)
    assert trial_decision(1.0, 0.5, 1.0, 0.5, 2.0, fixed) == (0.5, True, 0.5, 2.0, False)
    assert trial_decision(1.0, 1e-5, 1.0, 0.5, 2.0, fixed) == (1e-5, False, 1.0, 2.0, False)
    assert trial_decision(1.0, 1e-5, 1.0, 1.0, 2.0, fixed)[1:] == (False, 1.0, 2.0, True)  # at lambda_max
    assert trial_decision(1.0, -1.0, -2.0, 0.5, 2.0, fixed)[:2] == (-math.inf, False)  # predicted <= 0
    assert trial_decision(1.0, -0.5, 1.0, 0.5, 2.0, fixed)[1] is False  # actual <= 0
    assert trial_decision(math.nan, 0.9, 1.0, 0.5, 2.0, fixed)[1] is False  # nonfinite cost
    q, acc, lam, rej, stop = trial_decision(0.0, 0.0, 0.0, 0.25, 2.0, fixed, usable=False)
    assert math.isnan(q) and (acc, lam, rej, stop) == (False, 1.0, 2.0, False)

    adaptive = TrustRegion(lambda_min=1e-6, lambda_max=1.0, strategy="adaptive")
    assert trial_decision(1.0, 2.0, 2.0, 0.75, 8.0, adaptive) == (1.0, True, 0.25, 2.0, False)  # x 1/3
    assert trial_decision(1.0, 1.0, 2.0, 0.75, 8.0, adaptive) == (0.5, True, 0.75, 2.0, False)  # x 1
    assert trial_decision(1.0, 1.5, 2.0, 0.5, 8.0, adaptive) == (0.75, True, 0.4375, 2.0, False)  # x 7/8
    assert trial_decision(1.0, 4.0, 2.0, 0.75, 8.0, adaptive)[2] == 0.25  # the gain ratio counts up to 1
    assert trial_decision(1.0, 1e-5, 1.0, 1e-3, 2.0, adaptive)[2:] == (2e-3, 4.0, False)
    assert trial_decision(1.0, 1e-5, 1.0, 1e-3, 8.0, adaptive)[2:] == (8e-3, 16.0, False)
    capped = TrustRegion(lambda_min=0.5, lambda_max=1.0, lambda_initial=0.5, strategy="adaptive")
    assert trial_decision(1.0, 1e-5, 1.0, 0.5, 2.0, capped)[2:] == (1.0, 2.0, False)  # rej <= max/min = 2

    assert trial_decision(1.0, -3.0, 1.0, 0.0, 2.0, None)[1:] == (True, 0.0, 2.0, False)  # no trust region
    assert trial_decision(math.inf, 1.0, 1.0, 0.0, 2.0, None)[1:] == (False, 0.0, 2.0, True)
