"""Heel-effect spectra for tests, standing in for SpekPy's off-axis spectra (absent here).

Kramers' law at the source's kVp behind its aluminium filtration, and a self-filtration in the
anode whose path grows as the take-off angle shrinks toward the anode side: t(s) = t_heel
sin(a) / sin(a - atan(s / sid)), with a the anode angle and s the heel coordinate on the SID plane
(s > 0 on the anode side).  Aluminium attenuation stands in for the anode's.  On the anode side fewer
photons leave, with harder spectra; beyond the take-off angle 0 (atan(s / sid) >= a) nothing is
emitted, like the field parts beyond the anode cut-off.  `provider` has the signature and return
value of chatcarlo_amd.spectrum.heel_spectra_for_source, `central_fluence` those of
spectrum.fluence_per_mas (the model at s = 0)."""
from __future__ import annotations

import math
import warnings

import numpy as np


def _bins(src, sid, s, t_heel_mm):
    from chatcarlo_amd.materials import linear_mu
    kvp = float(src["kvp"])
    filt = float(src.get("filtration_mm_al", 2.5))
    a = math.radians(float(src.get("anode_angle_deg", 12.0)))
    e = np.arange(5.5, kvp, 1.0)
    mu = linear_mu("aluminum", e)
    base = np.clip(e * (kvp - e), 0.0, None) * np.exp(-mu * filt / 10.0)
    take_off = a - math.atan(s / sid)
    if take_off <= 0.0:
        return e, np.zeros_like(base)
    return e, base * np.exp(-mu * t_heel_mm * math.sin(a) / math.sin(take_off) / 10.0)


def make_provider(t_heel_mm: float = 4.0, n_bins: int = 15):
    def provider(src, sid, span_cm, n_bins_=n_bins):
        centres = (np.arange(n_bins_) + 0.5) * (2.0 * span_cm / n_bins_) - span_cm
        spectra, flu = [], np.zeros(n_bins_)
        for k, s in enumerate(centres):
            e, w = _bins(src, sid, float(s), t_heel_mm)
            flu[k] = w.sum()
            spectra.append((e, w / w.sum() if w.sum() > 0 else w))
        if not flu.max() > 0:
            raise ValueError("heel effect: no emission in any bin")
        if (flu <= 0).any():
            warnings.warn("the anode-side edge of the field lies beyond the anode cut-off (take-off angle 0): "
                          "no photons are emitted there; narrow the field or steepen the anode angle",
                          stacklevel=2)
        return centres, tuple(spectra), flu / flu.max(), flu

    return provider


def make_central_fluence(t_heel_mm: float = 4.0):
    def central(src, sid):
        return float(_bins(src, sid, 0.0, t_heel_mm)[1].sum())
    return central


def expected_coordinate_probabilities(edges, centres, rel_flu, lo, hi):
    """Probabilities of heel coordinates in the bins `edges` when s is uniform on [lo, hi] and
    kept with probability interp(s, centres, rel_flu): integrals of the interpolated fluence."""
    x = np.linspace(lo, hi, 200_001)
    f = np.interp(x, centres, rel_flu)
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (f[1:] + f[:-1]) * np.diff(x))])
    at = np.interp(np.clip(edges, lo, hi), x, cum)
    p = np.diff(at)
    return p / p.sum()
