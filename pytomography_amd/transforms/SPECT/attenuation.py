"""SPECT attenuation transform (transforms/SPECT/attenuation.py:23-124).

Inside ``SPECTSystemMatrix`` the attenuation factor
exp(-(sum_{i'>=i} mu_rot dx - mu_rot(i) dx / 2)) is computed on the fly by
the fused HIP kernels (rotation of mu, FP64 reverse cumsum, exp); this class
carries the attenuation map and validates it.
"""
from __future__ import annotations

import torch

import pytomography_amd
from ..transform import Transform


class SPECTAttenuationTransform(Transform):
    """obj2obj attenuation transform.

    Args:
        attenuation_map: tensor [Lx, Ly, Lz] of attenuation coefficients (1/cm)
            aligned with the SPECT object (attenuation.py:40-48).
        filepath: CT DICOM series -- not supported (reading DICOM needs pydicom,
            and the CT -> mu conversion lives outside the projector path).
        assume_padded: must be True (the reference default, attenuation.py:53):
            the kernels always work in the padded rotation frame.
    """

    def __init__(self, attenuation_map: torch.Tensor | None = None, filepath=None,
                 mode: str = "constant", assume_padded: bool = True,
                 HU2mu_technique="from_table") -> None:
        super().__init__()
        if attenuation_map is None and filepath is None:
            raise Exception("Please supply only one of `attenuation_map` or `filepath` as arguments")
        if filepath is not None:
            raise NotImplementedError(
                "CT DICOM -> attenuation map conversion is not supported by pytomography_amd; "
                "pass attenuation_map (1/cm) aligned with the object")
        if not assume_padded:
            raise NotImplementedError("assume_padded=False is not supported by the fused kernels")
        self.filepath = None
        self.mode = mode
        self.assume_padded = assume_padded
        self.HU2mu_technique = HU2mu_technique
        self.attenuation_map = attenuation_map.to(self.device)

    def configure(self, object_meta, proj_meta) -> None:
        super().configure(object_meta, proj_meta)
        if tuple(self.attenuation_map.shape) != tuple(object_meta.shape):
            raise ValueError(f"attenuation map shape {tuple(self.attenuation_map.shape)} does not "
                             f"match the object shape {tuple(object_meta.shape)}")

    def forward(self, object_i, ang_idx):
        raise NotImplementedError(
            "SPECTAttenuationTransform is applied inside SPECTSystemMatrix's fused kernels")

    backward = forward
