"""Interfile header access (io/shared/interfile.py:10-68)."""
from __future__ import annotations

import os
import re
from pathlib import Path

import numpy as np
import torch

import pytomography_amd


def _escape(header: str) -> str:
    # brackets and parentheses are literal; the rest is a regular expression,
    # as in interfile.py:28
    return header.replace('[', r'\[').replace(']', r'\]').replace('(', r'\(').replace(')', r'\)')


def get_header_value(list_of_attributes, header: str, dtype: type = np.float32, split_substr=':=',
                     split_idx=-1, return_all=False):
    """Value of the first line of an Interfile header (a sequence of lines)
    matching ``header``; False when no line matches (interfile.py:10-45)."""
    pattern = re.compile(_escape(header))
    lines = [line for line in list_of_attributes if pattern.search(line)]
    if len(lines) == 0:
        return False
    values = []
    for line in lines:
        field = line.replace('\n', '').split(split_substr)[split_idx]
        if dtype == np.float32:
            values.append(np.float32(field))
        elif dtype == str:
            values.append(field.replace(' ', ''))
        elif dtype == int:
            values.append(int(field.replace(' ', '')))
        if not return_all:
            return values[0]
    return values


def read_header(headerfile: str) -> list[str]:
    with open(headerfile) as f:
        return f.readlines()


def get_attenuation_map_interfile(headerfile: str) -> torch.Tensor:
    """float32 map [x, y, z] from a SIMIND/Interfile density file, with the
    first and second axes of the file order reversed (interfile.py:47-68)."""
    h = read_header(headerfile)
    shape = (get_header_value(h, 'matrix size [3]', int), get_header_value(h, 'matrix size [2]', int),
             get_header_value(h, 'matrix size [1]', int))
    imagefile = get_header_value(h, 'name of data file', str)
    amap = np.fromfile(os.path.join(str(Path(headerfile).parent), imagefile), dtype=np.float32)
    amap = np.transpose(amap.reshape(shape), (2, 1, 0))[:, ::-1, ::-1]
    return torch.tensor(amap.copy()).to(pytomography_amd.device)
