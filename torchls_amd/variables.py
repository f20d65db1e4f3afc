"""Variable parameterizations (the subset of torchls.variables the path needs).

A ``Var`` subclass declares its storage default, tangent dimension, retraction
and optional retraction JVP at zero, as in the reference (variables.py:18-65).
``EuclideanParameterization`` is the implicit manifold of a named parameter
group without a declared manifold (variables.py:82-97).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import torch
from torch import Tensor


class Var:
    """Base of manifold parameterizations, configured through subclass keywords."""

    def __init_subclass__(cls, *, default_factory, tangent_dim=None, retract_fn=None,
                          retract_jvp_fn=None, **kwargs):
        super().__init_subclass__(**kwargs)
        default = default_factory()
        if not isinstance(default, Tensor) or not default.is_floating_point():
            raise TypeError("default_factory must return a floating point Tensor")
        if (tangent_dim is None) != (retract_fn is None):
            raise ValueError("Specify tangent_dim and retract_fn together")
        if retract_jvp_fn is not None and retract_fn is None:
            raise ValueError("retract_jvp_fn requires a manifold retract_fn")
        cls.default_factory = staticmethod(default_factory)
        cls.is_euclidean = retract_fn is None
        cls.tangent_dim = default.numel() if tangent_dim is None else int(tangent_dim)
        if cls.tangent_dim <= 0:
            raise ValueError("tangent_dim must be positive")
        cls.retract_fn = staticmethod(
            (lambda x, d: x + d.reshape(x.shape)) if retract_fn is None else retract_fn
        )
        cls.retract_jvp_fn = staticmethod(retract_jvp_fn) if retract_jvp_fn is not None else None


@dataclass(frozen=True)
class EuclideanParameterization:
    shape: tuple

    is_euclidean = True
    retract_jvp_fn = None

    @property
    def tangent_dim(self) -> int:
        return math.prod(self.shape)

    def default_factory(self) -> Tensor:
        return torch.zeros(self.shape)

    @staticmethod
    def retract_fn(value: Tensor, delta: Tensor) -> Tensor:
        return value + delta.reshape(value.shape)
