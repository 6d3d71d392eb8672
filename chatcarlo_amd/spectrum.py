"""Source energy spectra: explicit list or the Kramers + Al fallback.

Follows chatcarlo/spectrum.py:55-82.  SpekPy spectra (chatcarlo/spectrum.py:36-52,
88-129) need the absent spekpy package and are not built; a scene that gives
`kvp` without `spectrum` gets the reference's own SpekPy-absent fallback,
Kramers' law E (kVp - E) on 60 bins from 5 keV to kVp attenuated by
exp(-mu_Al t).
"""
from __future__ import annotations

import warnings

import numpy as np

AL_DENSITY = 2.699   # chatcarlo/materials.py:49


def _mu_al(e_keV):
    from .materials import linear_mu
    return linear_mu("aluminum", e_keV)


def kramers_fallback_spectrum(kvp: float, filtration_mm_al: float, n_bins: int = 60):
    """chatcarlo/spectrum.py:55-62: bin centres and normalised weights."""
    e = np.linspace(5.0, kvp, n_bins + 1)
    mid = 0.5 * (e[:-1] + e[1:])
    raw = np.clip(mid * (kvp - mid), 0, None)
    w = raw * np.exp(-_mu_al(mid) * filtration_mm_al / 10.0)
    return mid, w / w.sum()


def kramers_spectrum_bins(kvp: float, filtration_mm_al: float, bin_keV: float = 1.0,
                          e_min: float = 5.0):
    """Kramers + Al on fixed-width bins, as an explicit `source.spectrum` list.

    Used to materialise the benchmark scenes' spectra (e.g. 80 kVp, 2.5 mm Al,
    1 keV bins) so every consumer reads identical bins.
    """
    edges = np.arange(e_min, kvp + 1e-9, bin_keV)
    mid = 0.5 * (edges[:-1] + edges[1:])
    raw = np.clip(mid * (kvp - mid), 0, None)
    w = raw * np.exp(-_mu_al(mid) * filtration_mm_al / 10.0)
    keep = w > 0
    w = w[keep] / w[keep].sum()
    return [{"energy_keV": float(e), "weight": float(x)} for e, x in zip(mid[keep], w)]


def spectrum_table(src: dict):
    """(energies keV, cumulative probabilities) of a source (chatcarlo/spectrum.py:73-82)."""
    spec = src.get("spectrum")
    if spec:
        e = np.array([s["energy_keV"] for s in spec], dtype=float)
        w = np.array([s["weight"] for s in spec], dtype=float)
        w = w / w.sum()
    else:
        warnings.warn("spekpy is not available: using the Kramers-law fallback spectrum",
                      stacklevel=2)
        e, w = kramers_fallback_spectrum(src["kvp"], src.get("filtration_mm_al", 2.5))
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    return e, cdf
