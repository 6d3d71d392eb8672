"""Source energy spectra: explicit list or the Kramers + Al fallback, and the heel-axis spectra.

Follows chatcarlo/spectrum.py:36-82.  A scene that gives `kvp` without `spectrum` gets SpekPy's
spectrum when spekpy is installed, else the reference's own SpekPy-absent fallback, Kramers' law
E (kVp - E) on 60 bins from 5 keV to kVp attenuated by exp(-mu_Al t), with a warning.  The heel
effect's off-axis spectra (chatcarlo/spectrum.py:88-152) and the mAs calibration's fluence come
from SpekPy when it is installed and raise otherwise, as the reference does; the GPU sampler
takes the heel spectra as data (per-bin CDFs and relative fluences, chatcarlo_amd/source.py).
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


_spectrum_cache: dict = {}


def spekpy_spectrum(kvp: float, filtration_mm_al: float, anode_angle_deg: float):
    """SpekPy's central-axis spectrum behind the Al filtration (chatcarlo/spectrum.py:36-52): bin
    energies and normalised weights, negative weights clipped to 0; None without spekpy."""
    key = (float(kvp), float(filtration_mm_al), float(anode_angle_deg))
    if key not in _spectrum_cache:
        try:
            import spekpy
        except ImportError:
            return None
        spk = spekpy.Spek(kvp=key[0], th=key[2])
        spk.filter("Al", key[1])
        e, phi = spk.get_spectrum(edges=False)
        w = np.clip(np.asarray(phi, dtype=float), 0.0, None)
        _spectrum_cache[key] = (np.asarray(e, dtype=float), w / w.sum())
    return _spectrum_cache[key]


def spectrum_table(src: dict):
    """(energies keV, cumulative probabilities) of a source (chatcarlo/spectrum.py:64-82): the
    given list, else SpekPy's spectrum for the kVp, else the Kramers fallback with a warning."""
    spec = src.get("spectrum")
    if spec:
        e = np.array([s["energy_keV"] for s in spec], dtype=float)
        w = np.array([s["weight"] for s in spec], dtype=float)
        w = w / w.sum()
    else:
        sp = spekpy_spectrum(src["kvp"], src.get("filtration_mm_al", 2.5), src.get("anode_angle_deg", 12.0))
        if sp is not None:
            e, w = sp
        else:
            warnings.warn("spekpy is not available: using the Kramers-law fallback spectrum",
                          stacklevel=2)
            e, w = kramers_fallback_spectrum(src["kvp"], src.get("filtration_mm_al", 2.5))
    # energies of zero weight are never drawn; dropping them keeps e.g. SpekPy's sub-keV bins, emptied
    # by the filtration, out of the sampler's [1, 150] keV check
    keep = w > 0
    e, w = e[keep], w[keep]
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    return e, cdf


def fluence_per_mas(src: dict, sid: float) -> float:
    """Photons per cm^2 per mAs on the central axis at distance sid (SpekPy's absolute fluence
    behind the source's Al filtration, chatcarlo/source.py:255-258).  Without spekpy this raises,
    as the reference does (the Kramers fallback has no absolute scale)."""
    try:
        import spekpy
    except ImportError:
        raise RuntimeError("photon-count calibration (mas) needs spekpy, which is not installed") from None
    spk = spekpy.Spek(kvp=float(src["kvp"]), th=float(src.get("anode_angle_deg", 12.0)), z=float(sid), mas=1.0)
    spk.filter("Al", float(src.get("filtration_mm_al", 2.5)))
    return float(spk.get_flu())


HEEL_N_BINS = 15   # bins along the heel axis (chatcarlo/spectrum.py:85)
_heel_cache: dict = {}


def heel_spectra_for_source(src: dict, sid: float, span_cm: float, n_bins: int = HEEL_N_BINS):
    """Spectra along the heel axis (chatcarlo/spectrum.py:88-152): bin centres s on the SID
    plane over [-span_cm, span_cm] (s > 0 on the anode side), each bin's (energies, normalised
    weights), the relative fluence (largest 1) and the fluence per mAs.  SpekPy's off-axis
    spectrum at lateral offset -s gives each bin (its x < 0 is the anode side).  Bins beyond the
    anode cut-off emit nothing; that is warned about, and a field with no emission raises.
    Without spekpy this raises RuntimeError, as the reference does."""
    kvp = float(src["kvp"])
    filt = float(src.get("filtration_mm_al", 2.5))
    angle = float(src.get("anode_angle_deg", 12.0))
    key = (kvp, filt, angle, float(sid), float(span_cm), int(n_bins))
    if key in _heel_cache:
        return _heel_cache[key]
    try:
        import spekpy
    except ImportError:
        raise RuntimeError("the heel effect needs spekpy, which is not installed") from None
    centres = (np.arange(n_bins) + 0.5) * (2.0 * span_cm / n_bins) - span_cm
    spectra, flu = [], np.zeros(n_bins)
    for k, s in enumerate(centres):
        spk = spekpy.Spek(kvp=kvp, th=angle, z=float(sid), x=-float(s))
        spk.filter("Al", filt)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")   # zero emission beyond the cut-off is reported below
            e, phi = spk.get_spectrum(edges=False)
            w = np.clip(np.asarray(phi, dtype=float), 0.0, None)
            if w.sum() > 0:
                flu[k] = float(spk.get_flu())
                w = w / w.sum()
        spectra.append((np.asarray(e, dtype=float), w))
    if not flu.max() > 0:
        raise ValueError("heel effect: no emission in any bin along the heel axis; check the anode angle "
                         "and the field")
    if (flu <= 0).any():
        warnings.warn("the anode-side edge of the field lies beyond the anode cut-off (take-off angle 0): "
                      "no photons are emitted there; narrow the field or steepen the anode angle", stacklevel=2)
    out = (centres, tuple(spectra), flu / flu.max(), flu)
    _heel_cache[key] = out
    return out
