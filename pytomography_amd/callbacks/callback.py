"""Callbacks run after each sub-iteration (callbacks/callback.py:5-62)."""
from __future__ import annotations

import torch


class Callback:
    def run(self, object: torch.Tensor, n_iter: int, n_subset: int) -> torch.Tensor:
        return object

    def finalize(self, object: torch.Tensor):
        return None


class MultiCallback(Callback):
    def __init__(self, callbacks: list[Callback]):
        self.callbacks = callbacks

    def run(self, object, n_iter, n_subset):
        for callback in self.callbacks:
            object = callback.run(object, n_iter, n_subset)
        return object

    def finalize(self, object):
        for callback in self.callbacks:
            callback.finalize(object)
