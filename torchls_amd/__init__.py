"""MI355X-native (gfx950) block-sparse normal-equation path of torchls.

Hand-written HIP kernels behind a C ABI (include/torchls_hip.h), driven from
PyTorch-ROCm host code that mirrors torchls's solver interface with
``backend="hip"``: ``Problem.solve()``, ``ConjugateGradientConfig``,
``pcg``/``block_jacobi``, and the plan/operator seams ``HipNormalPlan``,
``HipDiagonalPlan`` and ``hip_pcg`` (drop-ins for WarpNormalPlan,
WarpDiagonalPlan and warp_pcg).
"""

from .geometry import SE3Var, se3_exp, se3_log, se3_retract_jvp, skew, so3_exp, so3_log
from .hip_diagonal import HipBlockJacobi, HipDiagonalPlan, LinAlgError
from .hip_normal import HipNormalPlan
from .hip_pcg import hip_pcg
from .losses import HuberLoss, Loss
from .problem import Linearization, Problem
from .solvers import (
    ConjugateGradientConfig,
    PCGResult,
    SolveSummary,
    TerminationConfig,
    TrustRegionConfig,
    block_jacobi,
    pcg,
    solve,
)
from .sparse import BlockJacobian, JacobianGroup, VariableLayout
from .variables import Var

__all__ = [
    "BlockJacobian", "ConjugateGradientConfig", "HipBlockJacobi", "HipDiagonalPlan", "HipNormalPlan",
    "HuberLoss", "JacobianGroup", "LinAlgError", "Linearization", "Loss", "PCGResult", "Problem", "SE3Var",
    "SolveSummary", "TerminationConfig", "TrustRegionConfig", "Var", "VariableLayout", "block_jacobi",
    "hip_pcg", "pcg", "se3_exp", "se3_log", "se3_retract_jvp", "skew", "so3_exp", "so3_log", "solve",
]
__version__ = "0.1.0"
