from .preconditioned_gradient_ascent import (PreconditionedGradientAscentAlgorithm,
                                             LinearPreconditionedGradientAscentAlgorithm,
                                             OSEM, BSREM, MLEM)

__all__ = ["PreconditionedGradientAscentAlgorithm", "LinearPreconditionedGradientAscentAlgorithm",
           "OSEM", "BSREM", "MLEM"]
