"""Differentiable SPECT projector.

The reference runs every transform under ``torch.no_grad``
(transforms/SPECT/attenuation.py:72,94; transforms/SPECT/psf.py:199,216), so
it has no gradient of a projection with respect to the activity or the
attenuation map.  ``spect_project`` is ``H(mu) f`` as a
``torch.autograd.Function`` whose backward is the hand-written
vector-Jacobian product of the HIP kernels (``spect_forward_vjp``):

* d/df  : the exact transpose of the forward (rotation transpose written as
  a gather, PSF and attenuation are symmetric / diagonal);
* d/dmu : through the reverse cumulative sum and exp of the attenuation
  factor (utils/misc.py:18, attenuation.py:21) and the transpose of the
  rotation of mu.
"""
from __future__ import annotations

import torch

from pytomography_amd.projectors.SPECT import SPECTSystemMatrix


class SPECTProjection(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f, mu, system_matrix: SPECTSystemMatrix, subset_idx):
        ctx.system_matrix = system_matrix
        ctx.subset_idx = subset_idx
        ctx.save_for_backward(f, mu if mu is not None else torch.empty(0))
        ctx.has_mu = mu is not None
        return system_matrix.forward_with_mu(f.detach(), None if mu is None else mu.detach(), subset_idx)

    @staticmethod
    def backward(ctx, grad_out):
        f, mu = ctx.saved_tensors
        mu = mu if ctx.has_mu else None
        need_mu = ctx.has_mu and ctx.needs_input_grad[1]
        gf, gm = ctx.system_matrix.forward_vjp(grad_out.contiguous(), f, mu, ctx.subset_idx,
                                               need_mu=need_mu)
        return (gf if ctx.needs_input_grad[0] else None), (gm if need_mu else None), None, None


def spect_project(f: torch.Tensor, mu: torch.Tensor | None, system_matrix: SPECTSystemMatrix,
                  subset_idx: int | None = None) -> torch.Tensor:
    """H(mu) f, differentiable with respect to f and mu."""
    return SPECTProjection.apply(f, mu, system_matrix, subset_idx)


def poisson_log_likelihood(f, mu, system_matrix, projections, additive_term=None, subset_idx=None):
    """L(f, mu) = sum(g log(H(mu) f + s) - (H(mu) f + s)), the objective whose
    gradient PoissonLogLikelihood.compute_gradient (poisson_log_likelihood.py:16-36)
    evaluates for f."""
    hf = spect_project(f, mu, system_matrix, subset_idx)
    if additive_term is not None:
        hf = hf + additive_term
    return (projections * torch.log(hf) - hf).sum()
