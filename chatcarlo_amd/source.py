"""Source description for the GPU sampler (chatcarlo/source.py:20-210).

The reference samples focal spots, directions and energies in NumPy per batch;
here the host only resolves the scene's source section into a cc_source record
(focal spot, beam basis, field, CT rotation / helical shift, spectrum CDF) and
the source kernel samples every history from its own Philox stream.  Uniform
slots per history: angle, helical shift, field coordinate a, field coordinate
b, energy, tau.  The heel effect and mAs calibration need SpekPy (absent) and
raise, as the reference does without SpekPy (chatcarlo/source.py:222-223,
spectrum.py:100-102).
"""
from __future__ import annotations

import math

import numpy as np

from . import _native as N
from .spectrum import spectrum_table

_AXES = {"x": np.array([1.0, 0.0, 0.0]), "y": np.array([0.0, 1.0, 0.0]), "z": np.array([0.0, 0.0, 1.0])}
_FIELD = {"rect": 0, "cone": 1, "parallel": 2}   # CC_FIELD_RECT / CONE / PARALLEL


def beam_basis(d: np.ndarray):
    """Local basis (u, v) orthogonal to the beam axis (chatcarlo/source.py:28-35)."""
    if abs(d[2]) < 0.999:
        u = np.array([-d[1], d[0], 0.0])
    else:
        u = np.array([1.0, 0.0, 0.0])
    u = u / np.linalg.norm(u)
    return u, np.cross(d, u)


def cone_half_angle_rad(fld: dict) -> float:
    """chatcarlo/source.py:38-40."""
    return math.atan((fld["diameter_cm"] / 2.0) / fld["sid_cm"])


class SourceRecord:
    """A cc_source plus the arrays it points to (kept alive with it)."""

    def __init__(self, src: dict):
        if src.get("heel_effect"):
            raise RuntimeError("the heel effect needs spekpy, which is not installed")
        self.spec_e, self.spec_cdf = spectrum_table(src)
        self.spec_e = np.ascontiguousarray(self.spec_e, dtype=np.float64)
        self.spec_cdf = np.ascontiguousarray(self.spec_cdf, dtype=np.float64)
        d = np.asarray(src["direction"], dtype=float)
        d = d / np.linalg.norm(d)
        u, v = beam_basis(d)
        fld = src["field"]
        shape = fld.get("shape", "rect")
        s = N.CcSource()
        s.field = _FIELD[shape]
        s.n_spec = len(self.spec_e)
        s.pos[:] = [float(x) for x in src["position"]]
        s.dir[:] = d.tolist()
        s.u[:] = u.tolist()
        s.v[:] = v.tolist()
        s.sid = float(fld.get("sid_cm") or 0.0)
        if shape in ("rect", "parallel"):
            s.width, s.height = (float(x) for x in fld["size_cm"])
        if shape == "cone":
            s.cos_half = math.cos(cone_half_angle_rad(fld))
        rot = src.get("rotation")
        if rot is not None:
            s.iso[:] = [float(x) for x in rot["isocenter"]]
            s.axis[:] = _AXES[rot.get("axis", "z")].tolist()
            na = rot.get("n_angles")
            if na:
                s.rotation, s.n_angles = 2, int(na)
            else:
                s.rotation = 1
            s.scan_length = float(rot.get("scan_length_cm") or 0.0)
        s.spec_e = N.ptr(self.spec_e, N.C.c_double)
        s.spec_cdf = N.ptr(self.spec_cdf, N.C.c_double)
        self.struct = s


def sample_source_photons(src: dict, n: int, rng: np.random.Generator):
    """n source photons (pos, dir, energy as float64 arrays) drawn on the GPU
    (chatcarlo/source.py:67-210 sample_source_photons).  The Philox seed is drawn from `rng`;
    photon h uses the stream of history h, as in run_transport."""
    import ctypes as C

    from .transport import default_engine
    eng = default_engine()
    rec = SourceRecord(src)
    n = int(n)
    pos = np.zeros((n, 3), np.float32)
    dr = np.zeros((n, 3), np.float32)
    en = np.zeros(n, np.float32)
    if n:
        seed = int(rng.integers(0, 2 ** 63))
        N.check(eng.lib.cc_test_source(eng.ctx, C.byref(rec.struct), n, seed, 0, N.ptr(pos, C.c_float),
                                       N.ptr(dr, C.c_float), N.ptr(en, C.c_float)))
    return pos.astype(np.float64), dr.astype(np.float64), en.astype(np.float64)
