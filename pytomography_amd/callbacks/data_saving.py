"""DataStorageCallback (callbacks/data_saving.py:7-49)."""
from __future__ import annotations

import torch

from .callback import Callback


class DataStorageCallback(Callback):
    """Stores every iterate and ``likelihood.projections_predicted`` on the CPU."""

    def __init__(self, likelihood, object_initial: torch.Tensor) -> None:
        self.object_previous = torch.clone(object_initial)
        self.objects = []
        self.projections_predicted = []
        self.likelihood = likelihood

    def run(self, object, n_iter, n_subset):
        self.objects.append(self.object_previous.cpu())
        self.projections_predicted.append(self.likelihood.projections_predicted.cpu())
        self.object_previous = torch.clone(object)
        return object

    def finalize(self, object):
        self.objects.append(object.cpu())
        self.projections_predicted.append(None)
