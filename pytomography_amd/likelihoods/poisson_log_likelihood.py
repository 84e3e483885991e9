"""Poisson log-likelihood (likelihoods/poisson_log_likelihood.py:7-36)."""
from __future__ import annotations

import torch

import pytomography_amd
from pytomography_amd import _lib
from pytomography_amd.projectors.SPECT import SPECTSystemMatrix
from .likelihood import Likelihood


class PoissonLogLikelihood(Likelihood):
    """L(g|f) = sum g log(Hf + s) - (Hf + s); gradient H^T(g/(Hf+s)) - H^T 1."""

    def poisson_ratio(self, proj_subset, hf, additive_subset):
        """(pred, ratio) = (hf + s, g / (hf + s + delta)) in one kernel."""
        if isinstance(self.system_matrix, SPECTSystemMatrix) and hf.device.type == "cuda":
            dt = hf.dtype
            g = _lib.require_device_tensor(proj_subset, "projections", dt)
            s = _lib.require_device_tensor(additive_subset, "additive_term", dt)
            pred = torch.empty_like(hf)
            ratio = torch.empty_like(hf)
            _lib.check(_lib.lib().spect_poisson_ratio(
                _lib.dtype_code(dt), g.data_ptr(), hf.contiguous().data_ptr(), s.data_ptr(),
                pred.data_ptr(), ratio.data_ptr(), hf.numel(), float(pytomography_amd.delta),
                _lib.stream_ptr(hf.device)))
            return pred, ratio
        pred = hf + additive_subset
        return pred, proj_subset / (pred + pytomography_amd.delta)

    def compute_gradient(self, object: torch.Tensor, subset_idx: int | None = None,
                         norm_BP_subset_method: str = "subset_specific") -> torch.Tensor:
        """poisson_log_likelihood.py:16-36; sets ``projections_predicted``."""
        proj_subset = self._get_projection_subset(self.projections, subset_idx)
        additive_term_subset = self._get_projection_subset(self.additive_term, subset_idx)
        hf = self.system_matrix.forward(object, subset_idx)
        self.projections_predicted, ratio = self.poisson_ratio(proj_subset, hf, additive_term_subset)
        norm_BP = self._get_normBP(subset_idx)
        return self.system_matrix.backward(ratio, subset_idx) - norm_BP

    # ---------------------------------------------------------- second order
    # Operators of poisson_log_likelihood.py:38-114, built on the accelerated
    # forward / reference back projection.  FP is H_m f (the additive term is
    # not included, as in the reference); note the three different placements
    # of delta.
    def _fp(self, object, precomputed_forward_projection, subset_idx):
        if precomputed_forward_projection is None:
            return self.system_matrix.forward(object, subset_idx)
        return precomputed_forward_projection

    def compute_gradient_ff(self, object, precomputed_forward_projection=None, subset_idx=None):
        """x -> -H^T (g / (FP^2 + delta) * H x)  (poisson_log_likelihood.py:38-61)."""
        FP = self._fp(object, precomputed_forward_projection, subset_idx)
        g = self._get_projection_subset(self.projections, subset_idx)
        weight = g / (FP ** 2 + pytomography_amd.delta)

        def operator(x):
            return -self.system_matrix.backward(self.system_matrix.forward(x, subset_idx) * weight,
                                                subset_idx)
        return operator

    def compute_gradient_gf(self, object, precomputed_forward_projection=None, subset_idx=None):
        """x -> H x / (FP + delta)  (poisson_log_likelihood.py:63-85)."""
        FP = self._fp(object, precomputed_forward_projection, subset_idx)
        scale = 1 / (FP + pytomography_amd.delta)

        def operator(x):
            return self.system_matrix.forward(x, subset_idx) * scale
        return operator

    def compute_gradient_sf(self, object, precomputed_forward_projection=None, subset_idx=None):
        """x -> -H x * g / (FP + delta)^2  (poisson_log_likelihood.py:87-114)."""
        g = self._get_projection_subset(self.projections, subset_idx)
        FP = self._fp(object, precomputed_forward_projection, subset_idx)
        weight = g / (FP + pytomography_amd.delta) ** 2

        def operator(x):
            return -self.system_matrix.forward(x, subset_idx) * weight
        return operator
