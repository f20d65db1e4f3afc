"""SE(3) between factor of the pose-graph workload (CPU checks)."""

import numpy as np
import scipy.linalg
import torch

from torchls_amd import generators
from torchls_amd.generators import se3_inv_np
from torchls_amd.geometry import se3_exp, se3_inv, se3_log
from torchls_amd.residuals import between


def _vee(L):
    return np.array([L[0, 3], L[1, 3], L[2, 3], L[2, 1], L[0, 2], L[1, 0]])


def test_se3_log_is_the_matrix_logarithm():
    rng = np.random.default_rng(0)
    xi = rng.normal(0.0, 0.4, (16, 6))
    xi[0] = 0.0
    xi[1, 3:] = 1e-7  # Taylor branch
    T = se3_exp(torch.tensor(xi))
    got = se3_log(T).numpy()
    assert np.abs(got - xi).max() <= 1e-12
    for k in range(2, 16):
        assert np.allclose(_vee(scipy.linalg.logm(T[k].numpy()).real), got[k], rtol=0, atol=1e-12)
    assert torch.allclose(se3_inv(T) @ T, torch.eye(4, dtype=T.dtype).expand(16, 4, 4), atol=1e-14)


def test_between_residual_is_the_measurement_error():
    """At the true poses the residual is Log(Z^-1 T_i^-1 T_j) of the generator's noise."""
    pg = generators.pose_graph(300, seed=3, loop_fraction=0.05, loop_radius=3.0)
    assert pg.n_loop > 0
    Ti, Tj, Z = (torch.tensor(a) for a in (pg.poses_true[pg.i], pg.poses_true[pg.j], pg.measurements))
    r = torch.vmap(between)(Ti, Tj, Z).numpy()
    E = se3_inv_np(pg.measurements) @ se3_inv_np(pg.poses_true[pg.i]) @ pg.poses_true[pg.j]
    for k in (0, 7, pg.n_factors - 1):
        assert np.allclose(_vee(scipy.linalg.logm(E[k]).real), r[k], rtol=0, atol=1e-12)
    assert 0.0 < np.abs(r).max() < 0.1  # noise sigma 0.01 per coordinate
