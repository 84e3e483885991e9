"""Prior base class (priors/prior.py:7-57)."""
from __future__ import annotations

import abc

import torch

import pytomography_amd


class Prior:
    """V(f) in the log-posterior ln L - beta V; subclasses give dV/df."""

    @abc.abstractmethod
    def __init__(self, beta: float, obj2obj_transforms=None):
        self.beta = beta
        self.obj2obj_transforms = []
        self.device = pytomography_amd.device

    def set_object_meta(self, object_meta) -> None:
        self.object_meta = object_meta

    def set_beta_scale(self, factor: float) -> None:
        self.beta_scale_factor = factor

    def set_FOV_scale(self, FOV_scale: torch.Tensor) -> None:
        self.FOV_scale = FOV_scale

    def set_object(self, object: torch.Tensor) -> None:
        self.object = object.clone()

    @abc.abstractmethod
    def __call__(self):
        ...
