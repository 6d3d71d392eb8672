"""h*(10)/Phi from ICRP-74 (chatcarlo/dose_coefficients.py:26-42), host form.

The transport kernels evaluate the same log-log interpolation on the GPU; this
host function serves reporting and set-up.
"""
from __future__ import annotations

import numpy as np

from .datapack import load_hstar10


def h_star_10_per_fluence(energies_keV) -> np.ndarray:
    """[pSv cm^2], log-log between ICRP-74 points, clamped to 10 keV - 10 MeV."""
    e_tab, h_tab = load_hstar10()
    e = np.atleast_1d(np.asarray(energies_keV, dtype=float))
    ec = np.clip(e, e_tab[0], e_tab[-1])
    return np.exp(np.interp(np.log(ec), np.log(e_tab), np.log(h_tab)))
