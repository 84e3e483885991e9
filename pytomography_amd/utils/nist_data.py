"""Linear attenuation from NIST mass-attenuation tables (utils/nist_data.py:9-66).

The tables themselves are not shipped: pass the path of a NIST XCOM-format
file (first line density in g/cm^3, then rows of energy [MeV], mu/rho
[cm^2/g], mu_en/rho [cm^2/g])."""
from __future__ import annotations

import re
import warnings

import numpy as np
from scipy.optimize import curve_fit


def dual_sqrt_exponential(energy, c1, c2, d1, d2):
    """c1 exp(-d1 sqrt(E)) + c2 exp(-d2 sqrt(E)) (nist_data.py:9-25)."""
    return c1 * np.exp(-d1 * np.sqrt(energy)) + c2 * np.exp(-d2 * np.sqrt(energy))


def get_E_mu_data_from_datasheet(file: str):
    """Energies (keV) and linear attenuation (1/cm) between 50 and 550 keV
    (nist_data.py:27-46)."""
    with open(file) as f:
        lines = f.readlines()
    rho = float(lines[0])
    rows = [re.split(r'\s+', line)[-4:-1] for line in lines[1:]]
    E, mu_rho, _ = np.array(rows).astype(float).T
    E = E * 1000
    mu = mu_rho * rho
    idx = (E > 50) * (E < 550)
    return E[idx], mu[idx]


def get_mu_from_spectrum_interp(file: str, energy: float):
    """mu at ``energy`` from a fit of dual_sqrt_exponential to the table
    (nist_data.py:48-66)."""
    Edata, mudata = get_E_mu_data_from_datasheet(file)
    with warnings.catch_warnings(), np.errstate(all="ignore"):
        warnings.simplefilter("ignore")
        p = curve_fit(dual_sqrt_exponential, Edata, mudata)[0]
    return dual_sqrt_exponential(energy, *p)
