"""Transform base class (transforms/transform.py:8-38)."""
from __future__ import annotations

import abc

import pytomography_amd


class Transform(metaclass=abc.ABCMeta):
    """Object-to-object or projection-to-projection mapping configured by a
    system matrix with its object/projection metadata."""

    def __init__(self) -> None:
        self.device = pytomography_amd.device

    def configure(self, object_meta, proj_meta) -> None:
        self.object_meta = object_meta
        self.proj_meta = proj_meta

    @abc.abstractmethod
    def forward(self, x, *args):
        ...

    @abc.abstractmethod
    def backward(self, x, *args):
        ...
