from .likelihood import Likelihood
from .poisson_log_likelihood import PoissonLogLikelihood

__all__ = ["Likelihood", "PoissonLogLikelihood"]
