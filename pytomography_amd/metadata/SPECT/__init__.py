from .spect_metadata import SPECTObjectMeta, SPECTProjMeta, SPECTPSFMeta

__all__ = ["SPECTObjectMeta", "SPECTProjMeta", "SPECTPSFMeta"]
