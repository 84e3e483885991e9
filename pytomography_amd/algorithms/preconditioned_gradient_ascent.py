"""Preconditioned gradient ascent: OSEM, BSREM, MLEM
(algorithms/preconditioned_gradient_ascent.py:13-168,301-329,436-474,502-510).

f^{n+1} = f^n + C^n(f^n) [grad L(g_m | f^n) - grad V(f^n)], then values
<= ``addition_after_iteration`` are set to it.  For OSEM on a
``SPECTSystemMatrix`` with a ``PoissonLogLikelihood`` each sub-iteration is
one fused chain of kernels (forward, ratio, back projection + update); BSREM
uses the RDP stencil kernel and a fused update kernel.
"""
from __future__ import annotations

from typing import Callable

import torch

import pytomography_amd
from pytomography_amd import _lib
from pytomography_amd.likelihoods import PoissonLogLikelihood
from pytomography_amd.projectors.SPECT import SPECTSystemMatrix


class PreconditionedGradientAscentAlgorithm:
    def __init__(self, likelihood, prior=None, object_initial: torch.Tensor | None = None,
                 addition_after_iteration: float = 0, **kwargs) -> None:
        self.likelihood = likelihood
        if object_initial is None:
            self.object_prediction = self.likelihood.system_matrix._get_object_initial(pytomography_amd.device)
        else:
            self.object_prediction = object_initial.to(pytomography_amd.device).to(pytomography_amd.dtype)
        self.prior = prior
        if self.prior is not None:
            self.prior.set_object_meta(self.likelihood.system_matrix.object_meta)
            self.prior.set_FOV_scale(self.likelihood.system_matrix._get_prior_FOV_scale())
        self.objects_stored = []
        self.projections_predicted_stored = []
        self.addition_after_iteration = addition_after_iteration

    def _set_n_subsets(self, n_subsets: int) -> None:
        self.n_subsets = n_subsets
        self.likelihood._set_n_subsets(n_subsets)

    def _compute_preconditioner(self, object, n_iter, n_subset):
        raise NotImplementedError

    def _compute_callback(self, n_iter: int, n_subset: int) -> None:
        self.object_prediction = self.callback.run(self.object_prediction, n_iter, n_subset)

    def _fused_subiteration(self, n_iter: int, subset_idx) -> bool:
        """Hook for algorithms with a fused kernel path; returns False if none applies."""
        return False

    def __call__(self, n_iters: int, n_subsets: int = 1, n_subset_specific: int | None = None,
                 callback=None):
        """preconditioned_gradient_ascent.py:80-129."""
        self.callback = callback
        self.n_iters = n_iters
        self._set_n_subsets(n_subsets)
        for j in range(n_iters):
            for k in range(n_subsets):
                if n_subset_specific is not None and n_subset_specific != k:
                    continue
                subset_idx = None if n_subsets == 1 else k
                if not self._fused_subiteration(j, subset_idx):
                    if bool(self.prior):
                        self.prior.set_object(torch.clone(self.object_prediction).to(pytomography_amd.device))
                        self.prior.set_beta_scale(self.likelihood.system_matrix.get_weighting_subset(subset_idx))
                        self.prior_gradient = self.prior(derivative_order=1)
                    else:
                        self.prior_gradient = 0
                    likelihood_gradient = self.likelihood.compute_gradient(self.object_prediction, subset_idx)
                    preconditioner = self._compute_preconditioner(self.object_prediction, j, subset_idx)
                    self.object_prediction += preconditioner * (likelihood_gradient - self.prior_gradient)
                    self.object_prediction[self.object_prediction <= self.addition_after_iteration] = \
                        self.addition_after_iteration
                if self.callback is not None:
                    self._compute_callback(n_iter=j, n_subset=k)
        if self.callback is not None:
            self.callback.finalize(self.object_prediction)
        return self.object_prediction


class LinearPreconditionedGradientAscentAlgorithm(PreconditionedGradientAscentAlgorithm):
    """C^n(f^n) = D^n f^n (preconditioned_gradient_ascent.py:131-168)."""

    def _linear_preconditioner_factor(self, n_iter: int, n_subset):
        raise NotImplementedError

    def _compute_preconditioner(self, object, n_iter, n_subset):
        return object * self._linear_preconditioner_factor(n_iter, n_subset)


def _fused_ok(alg) -> bool:
    lk = alg.likelihood
    return (isinstance(lk, PoissonLogLikelihood) and isinstance(lk.system_matrix, SPECTSystemMatrix)
            and alg.object_prediction.device.type == "cuda"
            and alg.object_prediction.dtype in (torch.float32, torch.float64)
            and alg.object_prediction.is_contiguous())


class OSEM(LinearPreconditionedGradientAscentAlgorithm):
    """OSEM: D^n = 1/(H_m^T 1 + delta) (preconditioned_gradient_ascent.py:301-329)."""

    def __init__(self, likelihood, object_initial: torch.Tensor | None = None):
        super().__init__(likelihood=likelihood, object_initial=object_initial)

    def _linear_preconditioner_factor(self, n_iter: int, n_subset):
        return 1 / (self.likelihood._get_normBP(n_subset) + pytomography_amd.delta)

    def _fused_subiteration(self, n_iter: int, subset_idx) -> bool:
        if type(self)._linear_preconditioner_factor is not OSEM._linear_preconditioner_factor:
            return False
        if not _fused_ok(self):
            return False
        lk = self.likelihood
        sm = lk.system_matrix
        g = lk._get_projection_subset(lk.projections, subset_idx)
        s = lk._get_projection_subset(lk.additive_term, subset_idx)
        norm = lk._get_normBP(subset_idx)
        self.prior_gradient = 0
        lk.projections_predicted = sm.osem_step(self.object_prediction, g, s, norm, subset_idx,
                                                floor_value=self.addition_after_iteration)
        return True


class MLEM(OSEM):
    """OSEM with one subset (preconditioned_gradient_ascent.py:502-510)."""

    def __call__(self, n_iters, callback=None):
        return super().__call__(n_iters, n_subsets=1, callback=callback)


class BSREM(LinearPreconditionedGradientAscentAlgorithm):
    """BSREM: D^n = relax(n)/(w_m sum_m H_m^T 1 + delta) with the prior's
    beta scaled by w_m (preconditioned_gradient_ascent.py:436-474)."""

    def __init__(self, likelihood, object_initial: torch.Tensor | None = None, prior=None,
                 relaxation_sequence: Callable = lambda _: 1, addition_after_iteration=1e-4):
        self.relaxation_sequence = relaxation_sequence
        super().__init__(likelihood=likelihood, object_initial=object_initial, prior=prior,
                         addition_after_iteration=addition_after_iteration)

    def _linear_preconditioner_factor(self, n_iter: int, n_subset):
        relaxation_factor = self.relaxation_sequence(n_iter)
        norm_BP = self.likelihood._get_normBP(n_subset, return_sum=True)
        norm_BP_weight = self.likelihood.system_matrix.get_weighting_subset(n_subset)
        return relaxation_factor / (norm_BP_weight * norm_BP + pytomography_amd.delta)

    def _fused_subiteration(self, n_iter: int, subset_idx) -> bool:
        if not _fused_ok(self):
            return False
        from pytomography_amd.priors import RelativeDifferencePrior
        if self.prior is not None and not isinstance(self.prior, RelativeDifferencePrior):
            return False
        lk = self.likelihood
        sm = lk.system_matrix
        f = self.object_prediction
        w = sm.get_weighting_subset(subset_idx)
        prior_grad = None
        if self.prior is not None:
            self.prior.set_object(f)
            self.prior.set_beta_scale(w)
            prior_grad = self.prior(derivative_order=1)
        self.prior_gradient = 0 if prior_grad is None else prior_grad
        g = lk._get_projection_subset(lk.projections, subset_idx)
        s = lk._get_projection_subset(lk.additive_term, subset_idx)
        hf = sm.forward(f, subset_idx)
        lk.projections_predicted, ratio = lk.poisson_ratio(g, hf, s)
        b = sm.backward(ratio, subset_idx)
        norm = lk._get_normBP(subset_idx)
        if not hasattr(self, "_norm_sum") or self._norm_sum_key != id(lk.norm_BPs if subset_idx is not None else lk.norm_BP):
            self._norm_sum = lk._get_normBP(subset_idx, return_sum=True).contiguous()
            self._norm_sum_key = id(lk.norm_BPs if subset_idx is not None else lk.norm_BP)
        dt = f.dtype
        _lib.check(_lib.lib().spect_bsrem_update(
            _lib.dtype_code(dt), f.data_ptr(), b.data_ptr(), norm.contiguous().data_ptr(),
            self._norm_sum.data_ptr(), _lib.ptr(prior_grad), f.numel(),
            float(self.relaxation_sequence(n_iter)), float(w), float(pytomography_amd.delta),
            float(self.addition_after_iteration), _lib.stream_ptr(f.device)))
        return True
