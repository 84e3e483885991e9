from .metadata import ObjectMeta, ProjMeta

__all__ = ["ObjectMeta", "ProjMeta"]
