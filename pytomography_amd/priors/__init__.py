from .prior import Prior
from .nearest_neighbour import (NearestNeighbourPrior, RelativeDifferencePrior, NeighbourWeight,
                                EuclideanNeighbourWeight)

__all__ = ["Prior", "NearestNeighbourPrior", "RelativeDifferencePrior", "NeighbourWeight",
           "EuclideanNeighbourWeight"]
