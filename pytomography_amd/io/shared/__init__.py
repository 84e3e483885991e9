from .interfile import get_header_value, get_attenuation_map_interfile

__all__ = ["get_header_value", "get_attenuation_map_interfile"]
