"""System-matrix base class (projectors/system_matrix.py:10-115)."""
from __future__ import annotations

import abc
from copy import copy

import torch

import pytomography_amd


class SystemMatrix:
    """Linear operator H (forward) and its back projection (backward) with
    object-to-object and projection-to-projection transforms."""

    def __init__(self, object_meta, proj_meta, obj2obj_transforms=None, proj2proj_transforms=None) -> None:
        self.obj2obj_transforms = list(obj2obj_transforms or [])
        self.proj2proj_transforms = list(proj2proj_transforms or [])
        self.object_meta = object_meta
        self.proj_meta = proj_meta
        self.initialize_transforms()

    def __mul__(self, scalar):
        """system_matrix.py:32-44: scale forward and backward by ``scalar``."""
        old_forward = copy(self.forward)
        old_backward = copy(self.backward)

        def new_forward(*args, **kwargs):
            return old_forward(*args, **kwargs) * scalar

        def new_backward(*args, **kwargs):
            return old_backward(*args, **kwargs) * scalar

        self.forward = new_forward
        self.backward = new_backward
        return self

    def __rmul__(self, scalar):
        return self.__mul__(scalar)

    def initialize_transforms(self) -> None:
        for transform in self.obj2obj_transforms:
            transform.configure(self.object_meta, self.proj_meta)
        for transform in self.proj2proj_transforms:
            transform.configure(self.object_meta, self.proj_meta)

    def _get_object_initial(self, device=None):
        if device is None:
            device = pytomography_amd.device
        return torch.ones(self.object_meta.shape).to(device)

    def _get_prior_FOV_scale(self):
        return torch.ones(self.object_meta.shape).to(pytomography_amd.device)

    @abc.abstractmethod
    def forward(self, object, **kwargs):
        ...

    @abc.abstractmethod
    def backward(self, proj, **kwargs):
        ...
