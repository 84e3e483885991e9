from .preconditioned_gradient_ascent import (PreconditionedGradientAscentAlgorithm,
                                             LinearPreconditionedGradientAscentAlgorithm,
                                             OSEM, BSREM, MLEM, OSMAPOSL, RBIEM, RBIMAP)

__all__ = ["PreconditionedGradientAscentAlgorithm", "LinearPreconditionedGradientAscentAlgorithm",
           "OSEM", "BSREM", "MLEM", "OSMAPOSL", "RBIEM", "RBIMAP"]
