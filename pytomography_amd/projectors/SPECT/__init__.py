from .spect_system_matrix import SPECTSystemMatrix

__all__ = ["SPECTSystemMatrix"]
