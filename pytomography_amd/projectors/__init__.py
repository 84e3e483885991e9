from .system_matrix import SystemMatrix, ExtendedSystemMatrix

__all__ = ["SystemMatrix", "ExtendedSystemMatrix"]
