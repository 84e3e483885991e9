"""Pixel / angle subsampling of projections, metadata and maps (io/SPECT/shared.py:7-99)."""
from __future__ import annotations

import torch
from torch.nn.functional import avg_pool2d, avg_pool3d


def subsample_projections(projections: torch.Tensor, N_pixel: int, N_angle: int, N_angle_start: int = 0):
    """Sum over N_pixel x N_pixel pixel blocks, every N_angle-th angle."""
    result = avg_pool2d(projections.unsqueeze(0), N_pixel).squeeze() * N_pixel ** 2
    return result[N_angle_start::N_angle]


def subsample_metadata(object_meta, proj_meta, N_pixel: int = 1, N_angle: int = 1, N_angle_start: int = 0):
    object_meta.dr = tuple([N_pixel * d for d in object_meta.dr])
    object_meta.shape = tuple([int(s / N_pixel) for s in object_meta.shape])
    object_meta.compute_padded_shape()
    object_meta.dx *= N_pixel
    object_meta.dy *= N_pixel
    object_meta.dz *= N_pixel
    proj_meta.dr = tuple([N_pixel * d for d in proj_meta.dr])
    proj_meta.shape = (int(proj_meta.shape[0] / N_angle), int(proj_meta.shape[1] / N_pixel),
                       int(proj_meta.shape[2] / N_pixel))
    proj_meta.compute_padded_shape()
    proj_meta.radii = proj_meta.radii[N_angle_start::N_angle]
    proj_meta.angles = proj_meta.angles[N_angle_start::N_angle]
    proj_meta.num_projections = int(proj_meta.num_projections / N_angle)
    return object_meta, proj_meta


def subsample_projections_and_modify_metadata(object_meta, proj_meta, projections: torch.Tensor,
                                              N_pixel: int = 1, N_angle: int = 1, N_angle_start: int = 0):
    object_meta, proj_meta = subsample_metadata(object_meta, proj_meta, N_pixel, N_angle, N_angle_start)
    out = torch.stack([subsample_projections(p, N_pixel, N_angle, N_angle_start) for p in projections], dim=0)
    return object_meta, proj_meta, out


def subsample_amap(amap: torch.Tensor, N: int) -> torch.Tensor:
    """Mean over N x N x N blocks."""
    return avg_pool3d(amap.unsqueeze(0).unsqueeze(0), N).squeeze()
