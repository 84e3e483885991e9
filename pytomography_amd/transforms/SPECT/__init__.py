from .attenuation import SPECTAttenuationTransform
from .cutoff import CutOffTransform
from .psf import SPECTPSFTransform

__all__ = ["SPECTAttenuationTransform", "SPECTPSFTransform", "CutOffTransform"]
