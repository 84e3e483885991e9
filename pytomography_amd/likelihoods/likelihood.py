"""Likelihood base class (likelihoods/likelihood.py:9-122)."""
from __future__ import annotations

from typing import Callable

import torch

import pytomography_amd


class Likelihood:
    """Holds the system matrix, the measured projections and the additive
    term, and caches the subset sensitivities H_m^T 1."""

    def __init__(self, system_matrix, projections: torch.Tensor | None = None,
                 additive_term: torch.Tensor | None = None,
                 additive_term_variance_estimate: Callable | None = None) -> None:
        self.system_matrix = system_matrix
        if projections is None:
            self.projections = torch.tensor([1.0]).to(pytomography_amd.device)
        else:
            self.projections = projections
        self.FP = None
        if type(additive_term) is torch.Tensor:
            self.additive_term = additive_term.to(self.projections.device).to(pytomography_amd.dtype)
            self.exists_additive_term = True
        else:
            self.additive_term = torch.zeros(self.projections.shape).to(
                self.projections.device).to(pytomography_amd.dtype)
            self.exists_additive_term = False
        self.n_subsets_previous = -1
        self.additive_term_variance_estimate = additive_term_variance_estimate

    def _set_n_subsets(self, n_subsets: int) -> None:
        """likelihood.py:38-56: one H_m^T 1 per subset, recomputed only when
        the subset count changes."""
        self.n_subsets = n_subsets
        if n_subsets < 2:
            self.norm_BP = self.system_matrix.compute_normalization_factor()
        else:
            self.system_matrix.set_n_subsets(n_subsets)
            if self.n_subsets_previous != self.n_subsets:
                self.norm_BPs = [self.system_matrix.compute_normalization_factor(k)
                                 for k in range(self.n_subsets)]
        self.n_subsets_previous = n_subsets

    def _get_projection_subset(self, projections: torch.Tensor, subset_idx: int | None = None):
        """likelihood.py:58-72.  The measured projections and the additive term
        do not change between iterations, so each subset's contiguous copy is
        kept until the tensor is modified (its version counter moves)."""
        if subset_idx is None:
            return projections
        cache = self.__dict__.setdefault("_subset_cache", {})
        key = (id(projections), subset_idx)
        hit = cache.get(key)
        if hit is not None and hit[0] is projections and hit[1] == projections._version:
            return hit[2]
        sub = self.system_matrix.get_projection_subset(projections, subset_idx).contiguous()
        cache[key] = (projections, projections._version, sub)
        return sub

    def _get_normBP(self, subset_idx, return_sum: bool = False):
        """likelihood.py:74-91."""
        if subset_idx is None:
            return self.norm_BP
        if return_sum:
            return torch.stack(self.norm_BPs).sum(axis=0)
        return self.norm_BPs[subset_idx].to(pytomography_amd.device)

    def compute_gradient(self, *args, **kwargs):
        raise NotImplementedError("Compute gradient not implemented for this likelihood function")
