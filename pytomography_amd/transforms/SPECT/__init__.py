from .attenuation import SPECTAttenuationTransform
from .psf import SPECTPSFTransform

__all__ = ["SPECTAttenuationTransform", "SPECTPSFTransform"]
