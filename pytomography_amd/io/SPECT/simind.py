"""SIMIND output readers (io/SPECT/simind.py:17-233): metadata, projections,
energy windows, attenuation map and the collimator PSF model, returned as
the metadata and device tensors the accelerated projector takes."""
from __future__ import annotations

import os
from pathlib import Path
from typing import Sequence

import numpy as np
import torch

import pytomography_amd
from pytomography_amd.io.shared.interfile import get_header_value, read_header
from pytomography_amd.metadata.SPECT import SPECTObjectMeta, SPECTProjMeta, SPECTPSFMeta

relation_dict = {'unsignedinteger': 'int', 'shortfloat': 'float', 'int': 'int'}


def get_metadata(headerfile: str, distance: str = 'cm', corrfile: str | None = None):
    """(SPECTObjectMeta, SPECTProjMeta) of a SIMIND projection header
    (simind.py:17-56): pixel sizes in cm, angles from the start angle and
    extent (negated and wrapped for CW), radii from the header or a .cor file."""
    scale_factor = {'mm': 1 / 10, 'cm': 1}[distance]
    h = read_header(headerfile)
    proj_dim1 = get_header_value(h, 'matrix size [1]', int)
    proj_dim2 = get_header_value(h, 'matrix size [2]', int)
    dx = get_header_value(h, 'scaling factor (mm/pixel) [1]', np.float32) / 10
    dz = get_header_value(h, 'scaling factor (mm/pixel) [2]', np.float32) / 10
    dr = (dx, dx, dz)
    extent_of_rotation = get_header_value(h, 'extent of rotation', np.float32)
    number_of_projections = get_header_value(h, 'number of projections', int)
    start_angle = get_header_value(h, 'start angle', np.float32)
    direction = get_header_value(h, 'direction of rotation', str)
    angles = np.linspace(-start_angle, -start_angle + extent_of_rotation, number_of_projections,
                         endpoint=False)
    if direction == 'CW':
        angles = -angles % 360
    if corrfile is not None:
        radii = np.loadtxt(corrfile) * scale_factor
    else:
        radius = get_header_value(h, 'Radius', np.float32) * scale_factor
        radii = np.ones(len(angles)) * radius
    object_meta = SPECTObjectMeta(dr, (proj_dim1, proj_dim1, proj_dim2))
    proj_meta = SPECTProjMeta((proj_dim1, proj_dim2), (dx, dz), angles, radii)
    return object_meta, proj_meta


def _get_projections_from_single_file(headerfile: str) -> torch.Tensor:
    """[angles, r, z] projections: the file's [angle][z][r] images with z
    reversed, transposed (simind.py:58-82)."""
    h = read_header(headerfile)
    num_proj = get_header_value(h, 'total number of images', int)
    proj_dim1 = get_header_value(h, 'matrix size [1]', int)
    proj_dim2 = get_header_value(h, 'matrix size [2]', int)
    number_format = relation_dict[get_header_value(h, 'number format', str)]
    nbytes = get_header_value(h, 'number of bytes per pixel', int)
    imagefile = get_header_value(h, 'name of data file', str)
    dtype = {('int', 1): np.int8, ('int', 2): np.int16, ('int', 4): np.int32, ('int', 8): np.int64,
             ('float', 2): np.float16, ('float', 4): np.float32, ('float', 8): np.float64}[(number_format, nbytes)]
    p = np.fromfile(os.path.join(str(Path(headerfile).parent), imagefile), dtype=dtype)
    p = np.transpose(p.reshape((num_proj, proj_dim2, proj_dim1))[:, ::-1], (0, 2, 1))
    return torch.tensor(p.copy()).to(pytomography_amd.device)


def get_projections(headerfiles, weights=None) -> torch.Tensor:
    """One file, a list of files (energy windows, stacked) or a list of lists
    (windows of weighted organ files, summed) (simind.py:84-113)."""
    if type(headerfiles) is str:
        return _get_projections_from_single_file(headerfiles)
    if isinstance(headerfiles[0], Sequence) and not isinstance(headerfiles[0], str):
        windows = []
        for files in headerfiles:
            organs = []
            for j, hf in enumerate(files):
                pj = get_projections(hf)
                if weights is not None:
                    pj *= weights[j]
                organs.append(pj)
            windows.append(torch.sum(torch.stack(organs), dim=0))
        return torch.stack(windows).squeeze()
    return torch.stack([get_projections(hf) for hf in headerfiles]).squeeze()


def get_energy_window_bounds(headerfile: str):
    h = read_header(headerfile)
    return (get_header_value(h, 'energy window lower level', np.float32),
            get_header_value(h, 'energy window upper level', np.float32))


def get_energy_window_width(headerfile: str) -> float:
    lwr, upr = get_energy_window_bounds(headerfile)
    return upr - lwr


def combine_projection_data(headerfiles: Sequence[str], weights: Sequence[float]):
    """sum_i w_i * projections_i (simind.py:147-164)."""
    projections = 0
    for hf, w in zip(headerfiles, weights):
        projections += get_projections(hf) * w
    return projections


def get_attenuation_map(headerfile: str, smi_index_22: int = 3) -> torch.Tensor:
    """SIMIND density/mu map [x, y, z] (simind.py:166-193): index 22 = 3 (mu)
    reverses z, 5 (mu-castor) reverses y and z."""
    h = read_header(headerfile)
    shape = (get_header_value(h, 'matrix size [3]', int), get_header_value(h, 'matrix size [2]', int),
             get_header_value(h, 'matrix size [1]', int))
    imagefile = get_header_value(h, 'name of data file', str)
    CT = np.fromfile(os.path.join(str(Path(headerfile).parent), imagefile), dtype=np.float32)
    CT = np.transpose(CT.reshape(shape), (2, 1, 0))
    if smi_index_22 == 5:
        CT = CT[:, ::-1, ::-1]
    elif smi_index_22 == 3:
        CT = CT[:, :, ::-1]
    return torch.tensor(CT.copy()).to(pytomography_amd.device)


def _sigma_fit(r, a, b, c):
    return np.sqrt((a * r + b) ** 2 + c ** 2)


def get_psfmeta_from_header(headerfile: str, min_sigmas=3, lead_attenuation: float | None = None,
                            nist_lead_file: str | None = None) -> SPECTPSFMeta:
    """Collimator PSF sigma(d) = sqrt((a d + b)^2 + c^2) from the header's
    hole diameter, thickness, photon energy and intrinsic resolution
    (simind.py:195-233).  The lead attenuation at the photon energy is
    ``lead_attenuation`` (1/cm) or fitted from a NIST table: ``nist_lead_file``
    or, by default, the shipped ``data/NIST_attenuation_data/lead.csv``."""
    h = read_header(headerfile)
    to_cm = 0.1 if float(get_header_value(h, 'program version ', str)[1:2]) >= 8 else 1
    FWHM2sigma = 1 / (2 * np.sqrt(2 * np.log(2)))
    hole_diameter = get_header_value(h, 'Collimator hole diameter', np.float32) * to_cm
    hole_length = get_header_value(h, 'Collimator thickness', np.float32) * to_cm
    energy_keV = get_header_value(h, 'Photon Energy', np.float32)
    if lead_attenuation is None:
        from pytomography_amd.utils.nist_data import get_mu_from_spectrum_interp
        if nist_lead_file is None:   # the shipped NIST lead table, as the reference (:220)
            from pytomography_amd.io.SPECT.dicom import DATA_DIR
            nist_lead_file = os.path.join(DATA_DIR, "NIST_attenuation_data", "lead.csv")
        lead_attenuation = get_mu_from_spectrum_interp(nist_lead_file, energy_keV)
    collimator_slope = hole_diameter / (hole_length - (2 / lead_attenuation)) * 1 / (2 * np.sqrt(2 * np.log(2)))
    intrinsic = get_header_value(h, 'Intrinsic FWHM for the camera', np.float32)
    if intrinsic is False:
        intrinsic_resolution = 0
    else:
        intrinsic_resolution = intrinsic * FWHM2sigma * to_cm * (140 / energy_keV) ** 0.5
    collimator_intercept = hole_diameter * FWHM2sigma
    return SPECTPSFMeta(sigma_fit_params=[collimator_slope, collimator_intercept, intrinsic_resolution],
                        sigma_fit=_sigma_fit, min_sigmas=min_sigmas)
