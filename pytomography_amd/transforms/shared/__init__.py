from .kem import KEMTransform
from .motion import DVFMotionTransform

__all__ = ["KEMTransform", "DVFMotionTransform"]
