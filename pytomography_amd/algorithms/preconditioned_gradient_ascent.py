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
    """C^n(f^n) = D^n f^n (preconditioned_gradient_ascent.py:131-168), and the
    uncertainty of a masked sum of the reconstruction propagated back through
    the stored iterates (:170-299)."""

    def _linear_preconditioner_factor(self, n_iter: int, n_subset):
        raise NotImplementedError

    def _compute_preconditioner(self, object, n_iter, n_subset):
        return object * self._linear_preconditioner_factor(n_iter, n_subset)

    def compute_uncertainty(self, mask: torch.Tensor, data_storage_callback,
                            subiteration_number: int | None = None, return_pct: bool = False,
                            include_additive_term: bool = False, post_recon_filter=None):
        """Absolute (and optionally percent) uncertainty of the sum of the
        reconstruction over ``mask`` (preconditioned_gradient_ascent.py:170-214)."""
        if subiteration_number is None:
            subiteration_number = len(data_storage_callback.objects) - 1
        final_recon = data_storage_callback.objects[subiteration_number].to(pytomography_amd.device)
        if post_recon_filter is not None:
            q = post_recon_filter(mask)
            final_recon = post_recon_filter(final_recon)
        else:
            q = mask.clone()
        V = 0
        for n in range(subiteration_number, 0, -1):
            V += self._compute_B(q, data_storage_callback, n - 1, include_additive_term=include_additive_term)
            if n > 1:
                q = self._compute_Q(q, data_storage_callback, n - 1)
        abs2 = torch.sum(V[0] * self.likelihood.projections * V[0])
        if include_additive_term:
            abs2 += torch.sum(self.likelihood.additive_term_variance_estimate(V[1]))
        uncertainty_abs = torch.sqrt(abs2).item()
        if not return_pct:
            return uncertainty_abs
        return uncertainty_abs, uncertainty_abs / (final_recon * mask).sum().item() * 100

    def _subset(self, n):
        return None if self.n_subsets == 1 else n % self.n_subsets

    def _compute_Q(self, input: torch.Tensor, data_storage_callback, n: int) -> torch.Tensor:
        """Q operator of the uncertainty recursion (:216-247)."""
        subset_idx = self._subset(n)
        cur = data_storage_callback.objects[n].to(pytomography_amd.device)
        nxt = data_storage_callback.objects[n + 1].to(pytomography_amd.device)
        fp = data_storage_callback.projections_predicted[n].to(pytomography_amd.device)
        grad_ff = self.likelihood.compute_gradient_ff(cur, fp, subset_idx)
        output = input * cur * self._linear_preconditioner_factor(None, subset_idx)
        if self.prior is not None:
            self.prior.set_beta_scale(self.likelihood.system_matrix.get_weighting_subset(subset_idx))
            self.prior.set_object(cur)
            output = grad_ff(output) - self.prior(derivative_order=2)(output)
        else:
            output = grad_ff(output)
        output += (nxt / (cur + pytomography_amd.delta)) * input
        return output

    def _compute_B(self, input: torch.Tensor, data_storage_callback, n: int,
                   include_additive_term: bool = False) -> torch.Tensor:
        """B operator of the uncertainty recursion (:249-299): projections of
        every angle, zero outside the subset."""
        sm = self.likelihood.system_matrix
        subset_idx = self._subset(n)
        if subset_idx is None:
            rows = torch.arange(sm.proj_meta.shape[0]).to(torch.long).to(pytomography_amd.device)
        else:
            rows = sm.subset_indices_array[subset_idx]
        cur = data_storage_callback.objects[n].to(pytomography_amd.device)
        fp = data_storage_callback.projections_predicted[n].to(pytomography_amd.device)
        output = input * cur * self._linear_preconditioner_factor(None, subset_idx)
        primary = self.likelihood.compute_gradient_gf(cur, fp, subset_idx)(output)
        shape = (*fp.shape[:-3], *sm.proj_meta.shape[-3:])
        total = torch.zeros((2 if include_additive_term else 1, *shape)).to(pytomography_amd.device)
        total[0, ..., rows, :, :] = primary
        if include_additive_term:
            total[1, ..., rows, :, :] = self.likelihood.compute_gradient_sf(cur, fp, subset_idx)(output)
        return total


def _fused_ok(alg) -> bool:
    lk = alg.likelihood
    return (isinstance(lk, PoissonLogLikelihood) and isinstance(lk.system_matrix, SPECTSystemMatrix)
            and lk.system_matrix.fused
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


class OSMAPOSL(PreconditionedGradientAscentAlgorithm):
    """One-step-late MAP-EM: C^n = f^n / (H_m^T 1 + grad V(f^n))
    (preconditioned_gradient_ascent.py:331-362)."""

    def __init__(self, likelihood, object_initial: torch.Tensor | None = None, prior=None):
        super().__init__(likelihood=likelihood, object_initial=object_initial, prior=prior)

    def _compute_preconditioner(self, object, n_iter, n_subset):
        return object / (self.likelihood._get_normBP(n_subset) + self.prior_gradient
                         + pytomography_amd.delta)


class RBIEM(LinearPreconditionedGradientAscentAlgorithm):
    """Rescaled block-iterative EM: C^n = f^n / (sum_m n_m * r_m + delta),
    r_m = max(n_m / (sum_m n_m + delta)) (preconditioned_gradient_ascent.py:364-398)."""

    def __init__(self, likelihood, object_initial: torch.Tensor | None = None, prior=None):
        super().__init__(likelihood=likelihood, object_initial=object_initial, prior=prior)

    def _compute_preconditioner(self, object, n_iter, n_subset):
        norm = self.likelihood._get_normBP(n_subset)
        total = self.likelihood._get_normBP(n_subset, return_sum=True)
        rm = torch.max(norm / (total + pytomography_amd.delta))
        return object / (total * rm + pytomography_amd.delta)


class RBIMAP(PreconditionedGradientAscentAlgorithm):
    """Rescaled block-iterative MAP (preconditioned_gradient_ascent.py:400-434)."""

    def __init__(self, likelihood, object_initial: torch.Tensor | None = None, prior=None):
        super().__init__(likelihood=likelihood, object_initial=object_initial, prior=prior)

    def _compute_preconditioner(self, object, n_iter, n_subset):
        norm = self.likelihood._get_normBP(n_subset)
        total = self.likelihood._get_normBP(n_subset, return_sum=True)
        v = self.prior_gradient
        rm = torch.max((norm + v) / (total + v + pytomography_amd.delta))
        return object / (total * rm + v + pytomography_amd.delta)
