from .callback import Callback, MultiCallback
from .data_saving import DataStorageCallback

__all__ = ["Callback", "MultiCallback", "DataStorageCallback"]
