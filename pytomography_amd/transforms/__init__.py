from .transform import Transform

__all__ = ["Transform"]
