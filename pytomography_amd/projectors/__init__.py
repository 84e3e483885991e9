from .system_matrix import SystemMatrix

__all__ = ["SystemMatrix"]
