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


class ExtendedSystemMatrix(SystemMatrix):
    """H' = sum_n v_n (x) B_n H_n A_n: projections gain a leading dimension of
    size N, one entry per system matrix (system_matrix.py:117-221)."""

    def __init__(self, system_matrices, obj2obj_transforms=None, proj2proj_transforms=None) -> None:
        self.object_meta = system_matrices[0].object_meta
        self.proj_meta = system_matrices[0].proj_meta
        self.system_matrices = system_matrices
        self.obj2obj_transforms = obj2obj_transforms
        self.proj2proj_transforms = proj2proj_transforms
        for i in range(len(self.system_matrices)):
            for ts in (self.obj2obj_transforms, self.proj2proj_transforms):
                if ts is not None and ts[i] is not None:
                    ts[i].configure(self.system_matrices[i].object_meta, self.system_matrices[i].proj_meta)

    def forward(self, object, subset_idx: int | None = None):
        projs = []
        for i in range(len(self.system_matrices)):
            object_i = object
            if self.obj2obj_transforms is not None and self.obj2obj_transforms[i] is not None:
                object_i = self.obj2obj_transforms[i].forward(object)
            proj_i = self.system_matrices[i].forward(object_i, subset_idx)
            if self.proj2proj_transforms is not None and self.proj2proj_transforms[i] is not None:
                proj_i = self.proj2proj_transforms[i].forward(proj_i)
            projs.append(proj_i)
        return torch.stack(projs)

    def backward(self, proj, subset_idx: int | None = None):
        objects = []
        for i in range(len(self.system_matrices)):
            proj_i = proj[i]
            if self.proj2proj_transforms is not None and self.proj2proj_transforms[i] is not None:
                proj_i = self.proj2proj_transforms[i].backward(proj_i)
            object_i = self.system_matrices[i].backward(proj_i, subset_idx)
            if self.obj2obj_transforms is not None and self.obj2obj_transforms[i] is not None:
                object_i = self.obj2obj_transforms[i].backward(object_i)
            objects.append(object_i)
        return torch.stack(objects).sum(axis=0)

    def set_n_subsets(self, n_subsets: int):
        for system_matrix in self.system_matrices:
            system_matrix.set_n_subsets(n_subsets)
        self.subset_indices_array = self.system_matrices[0].subset_indices_array

    def get_projection_subset(self, projections, subset_idx):
        return self.system_matrices[0].get_projection_subset(projections, subset_idx)

    def get_weighting_subset(self, subset_idx):
        return self.system_matrices[0].get_weighting_subset(subset_idx)

    def compute_normalization_factor(self, subset_idx: int | None = None):
        """H'^T 1 over the subset's angles (system_matrix.py:209-221)."""
        norm_proj = torch.ones((len(self.system_matrices), *self.proj_meta.shape)).to(pytomography_amd.device)
        if subset_idx is not None:
            norm_proj = norm_proj[:, self.subset_indices_array[subset_idx]]
        return self.backward(norm_proj, subset_idx)
