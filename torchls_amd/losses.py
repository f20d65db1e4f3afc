"""Block robust losses on each observation's squared residual norm s = ||r_i||^2.

Same contract as torchls.losses (losses.py:13-107): ``evaluate(s)`` returns
(rho, rho', rho''); the Gauss-Newton model scales residual and Jacobian rows by
sqrt(rho'(s)), with Ceres's clamped curvature correction for a positive rho''
(the general ``Loss._linearize``); ``HuberLoss`` drops the correction and uses a
divided difference for trial reductions.  The float64 objective and trial
reductions of ``HuberLoss`` and of the plain squared norm run in HIP
(tlh_trial_objective); other Loss subclasses are evaluated with their own
tensor code.
"""

from __future__ import annotations

import math
from abc import ABC, abstractmethod
from dataclasses import dataclass

import torch
from torch import Tensor


class Loss(ABC):
    @abstractmethod
    def evaluate(self, squared_norm: Tensor) -> tuple[Tensor, Tensor, Tensor]:
        """Return rho(s), rho'(s), rho''(s)."""

    def value(self, squared_norm: Tensor) -> Tensor:
        return self.evaluate(squared_norm)[0]

    def _linearize(self, residual, blocks, squared_norm, evaluation):
        _, first, second = evaluation
        tiny = torch.finfo(squared_norm.dtype).tiny
        curved = (squared_norm > 0) & (second > 0)
        ratio = torch.where(curved, 2 * squared_norm * second / first.clamp_min(tiny), 0)
        alpha = -ratio / (1 + (1 + ratio).sqrt())
        correction = (alpha / squared_norm.clamp_min(tiny)).to(residual.dtype)
        weight = first.sqrt().to(residual.dtype)
        out = []
        for block in blocks:
            projected = torch.einsum("nr,nrd->nd", residual, block)
            out.append(weight[:, None, None] * (block - correction[:, None, None] * residual[:, :, None]
                                                * projected[:, None, :]))
        model = residual * (weight / (1 - alpha).to(residual.dtype))[:, None]
        return model, tuple(out)

    def _reduction(self, before, after, squared_difference):
        return self.value(before) - self.value(after)


@dataclass(frozen=True)
class HuberLoss(Loss):
    """rho(s) = s for sqrt(s) <= delta, else 2 delta sqrt(s) - delta^2 (block norm)."""

    delta: float = 1.0

    def __post_init__(self):
        if not math.isfinite(self.delta) or self.delta <= 0:
            raise ValueError("Huber delta must be finite and positive")

    def evaluate(self, squared_norm: Tensor):
        tiny = torch.finfo(squared_norm.dtype).tiny
        s = squared_norm.clamp_min(tiny)
        norm = s.sqrt()
        outer = squared_norm > self.delta * self.delta
        first = torch.where(outer, self.delta / norm, torch.ones_like(norm))
        value = torch.where(outer, self.delta * (2 * norm - self.delta), squared_norm)
        second = torch.where(outer, -0.5 * first / s, torch.zeros_like(norm))
        return value, first, second

    def value(self, squared_norm: Tensor) -> Tensor:
        norm = squared_norm.clamp_min(torch.finfo(squared_norm.dtype).tiny).sqrt()
        return torch.where(squared_norm > self.delta * self.delta, self.delta * (2 * norm - self.delta),
                           squared_norm)

    def _linearize(self, residual, blocks, squared_norm, evaluation):
        w = evaluation[1].sqrt().to(residual.dtype)
        return residual * w[:, None], tuple(b * w[:, None, None] for b in blocks)

    def _reduction(self, before, after, squared_difference):
        a, b = before.sqrt(), after.sqrt()
        low, high = torch.minimum(a, b), torch.maximum(a, b)
        tiny = torch.finfo(before.dtype).tiny
        shortfall = (self.delta - low).clamp_min(0)
        crossing = (2 * self.delta - shortfall * (shortfall / (high - low).clamp_min(tiny))) / (a + b).clamp_min(tiny)
        coef = torch.where(high <= self.delta, torch.ones_like(before),
                           torch.where(low >= self.delta, 2 * self.delta / (a + b).clamp_min(tiny), crossing))
        return squared_difference * coef
