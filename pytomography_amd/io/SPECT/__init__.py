from . import dicom, simind
from .shared import (subsample_projections, subsample_metadata,
                     subsample_projections_and_modify_metadata, subsample_amap)

__all__ = ["dicom", "simind", "subsample_projections", "subsample_metadata",
           "subsample_projections_and_modify_metadata", "subsample_amap"]
