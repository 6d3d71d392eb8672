"""A heel-effect spectra provider for tests, standing in for SpekPy's off-axis spectra (absent
here).  Kramers' law at the source's kVp behind its aluminium filtration and an anode
self-filtration that grows toward the anode side, t(s) = t_heel (1 + s / span): on the anode side
(s > 0) fewer photons leave, with a harder spectrum, as in the reference's heel effect.  Bins
beyond `cutoff_cm` (if given) emit nothing, like field parts beyond the anode cut-off.  Same
signature and return value as chatcarlo_amd.spectrum.heel_spectra_for_source."""
from __future__ import annotations

import warnings

import numpy as np


def make_provider(t_heel_mm: float = 4.0, cutoff_cm: float | None = None, n_bins: int = 15):
    from chatcarlo_amd.materials import linear_mu

    def provider(src, sid, span_cm, n_bins_=n_bins):
        kvp = float(src["kvp"])
        filt = float(src.get("filtration_mm_al", 2.5))
        e = np.arange(5.5, kvp, 1.0)
        mu = linear_mu("aluminum", e)
        base = np.clip(e * (kvp - e), 0.0, None) * np.exp(-mu * filt / 10.0)
        centres = (np.arange(n_bins_) + 0.5) * (2.0 * span_cm / n_bins_) - span_cm
        spectra, flu = [], np.zeros(n_bins_)
        for k, s in enumerate(centres):
            w = base * np.exp(-mu * t_heel_mm * (1.0 + s / span_cm) / 10.0)
            if cutoff_cm is not None and s > cutoff_cm:
                w = np.zeros_like(w)
            flu[k] = w.sum()
            spectra.append((e.copy(), w / w.sum() if w.sum() > 0 else w))
        if (flu <= 0).any():
            warnings.warn("the anode-side edge of the field lies beyond the anode cut-off (take-off angle 0): "
                          "no photons are emitted there; narrow the field or steepen the anode angle",
                          stacklevel=2)
        return centres, tuple(spectra), flu / flu.max(), flu

    return provider


def expected_coordinate_probabilities(edges, centres, rel_flu, lo, hi):
    """Probabilities of heel coordinates in the bins `edges` when s is uniform on [lo, hi] and
    kept with probability interp(s, centres, rel_flu): integrals of the interpolated fluence."""
    x = np.linspace(lo, hi, 200_001)
    f = np.interp(x, centres, rel_flu)
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (f[1:] + f[:-1]) * np.diff(x))])
    at = np.interp(np.clip(edges, lo, hi), x, cum)
    p = np.diff(at)
    return p / p.sum()
