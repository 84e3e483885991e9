from .kem_system_matrix import KEMSystemMatrix
from .motion_correction_system_matrix import MotionSystemMatrix

__all__ = ["KEMSystemMatrix", "MotionSystemMatrix"]
