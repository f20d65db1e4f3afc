"""HIP SE(3) retraction T Exp(delta) vs the matrix exponential of the twist."""

import numpy as np
import pytest
import scipy.linalg
import torch

from torchls_amd import Problem, SE3Var, generators
from torchls_amd.residuals import reprojection

pytestmark = pytest.mark.gpu


def twist(xi):
    v, w = xi[:3], xi[3:]
    m = np.zeros((4, 4))
    m[:3, :3] = [[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]]
    m[:3, 3] = v
    return m


@pytest.mark.parametrize("rows", [4, 3])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 2e-6)])
def test_se3_retract_matches_expm(device, rows, dtype, tol):
    print(f"se3 retract rows={rows} dtype={dtype}", flush=True)
    prob = generators.ba_small()
    rng = np.random.default_rng(4)
    n = prob.n_cameras
    # small angles (series branch), moderate and large angles
    scale = np.concatenate((np.full(n // 3, 1e-3), np.full(n // 3, 0.3), np.full(n - 2 * (n // 3), 2.5)))
    xi = rng.standard_normal((n, 6)) * scale[:, None]
    poses = prob.cameras_init
    P = Problem({"cams": torch.tensor(poses[:, :rows], dtype=dtype, device=device),
                 "points": torch.tensor(prob.points_init, dtype=dtype, device=device)}, manifolds={"cams": SE3Var})
    P.add_residual(reprojection, ("cams", torch.tensor(prob.cam_idx)), ("points", torch.tensor(prob.pt_idx)),
                   torch.tensor(prob.observations, dtype=dtype, device=device))
    an = P._analyze()
    delta = np.concatenate((xi.reshape(-1), rng.standard_normal(3 * prob.n_points) * 1e-2))
    d = torch.tensor(delta, dtype=torch.float64, device=device)
    out = an.retract(P.parameters, d, kernel=True)
    ref_t = an.retract(P.parameters, d, kernel=False)
    base = P.parameters["cams"].double().cpu().numpy()
    full = np.concatenate((base, np.tile([[[0.0, 0, 0, 1]]], (n, 1, 1))), axis=1) if rows == 3 else base
    ref = np.stack([full[i] @ scipy.linalg.expm(twist(xi[i])) for i in range(n)])[:, :rows]
    got = out["cams"].double().cpu().numpy()
    assert out["cams"].dtype == dtype and out["cams"].shape == (n, rows, 4)
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()
    assert np.abs(got - ref_t["cams"].double().cpu().numpy()).max() <= tol * np.abs(ref).max()
    pts = P.parameters["points"].double().cpu().numpy() + torch.tensor(delta[6 * n:], dtype=dtype).double().numpy().reshape(-1, 3)
    assert np.abs(out["points"].double().cpu().numpy() - pts).max() <= tol
