"""Source description for the GPU sampler (chatcarlo/source.py:20-210).

The reference samples focal spots, directions and energies in NumPy per batch;
here the host only resolves the scene's source section into a cc_source record
(focal spot, beam basis, field, CT rotation / helical shift, spectrum CDF) and
the source kernel samples every history from its own Philox stream.  Uniform
slots per history: angle, helical shift, field coordinate a, field coordinate
b, energy, tau; with the heel effect, the keep decisions and the field points
that replace rejected ones come from the history's separate heel stream.  The
heel effect's spectra (spectrum.heel_spectra_for_source) and the mAs
calibration's fluence (spectrum.fluence_per_mas) come from SpekPy, which is
absent here: both raise, as the reference does without SpekPy
(chatcarlo/spectrum.py:100-102, source.py:222-223).
"""
from __future__ import annotations

import math

import numpy as np

from . import _native as N
from . import spectrum as _spectrum
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


def heel_axis_coeffs(src: dict, d, u, v):
    """The anode direction projected on the plane across the beam, as coefficients of the beam
    basis (u, v), so that the heel axis turns with a rotating tube (chatcarlo/source.py:43-56)."""
    a = np.asarray(src["anode_direction"], dtype=float)
    a = a - (a @ d) * d
    norm = np.linalg.norm(a)
    if norm < 1e-9:
        raise ValueError("source.anode_direction is parallel to the beam: the anode-cathode axis needs a "
                         "component across the central axis")
    a = a / norm
    return float(a @ u), float(a @ v)


def heel_span_cm(fld: dict, hu: float, hv: float) -> float:
    """Half-width of the field along the heel axis on the SID plane (chatcarlo/source.py:59-64)."""
    if fld.get("shape", "rect") == "cone":
        return fld["diameter_cm"] / 2.0
    w, h = fld["size_cm"]
    return abs(hu) * w / 2.0 + abs(hv) * h / 2.0


def heel_tables(centres, spectra, rel_flu):
    """The heel spectra as sampler tables: one energy grid (every energy some bin emits at),
    each bin's CDF on it, and the first and last bin of non-zero fluence.  A bin that emits
    nothing (beyond the anode cut-off) takes the CDF of the nearest bin that does; the sampler
    never draws from it.  Returns (energies, cdf rows, relative fluences, lo, hi)."""
    rel = np.asarray(rel_flu, dtype=float)
    nz = np.flatnonzero(rel > 0)
    if len(nz) == 0:
        raise ValueError("heel effect: zero fluence in every bin")
    grid = np.unique(np.concatenate([np.asarray(e, dtype=float)[np.asarray(w) > 0] for e, w in spectra]))
    if grid[0] < 1.0 or grid[-1] > 150.0:
        raise ValueError(f"heel spectra emit at {grid[0]:g}-{grid[-1]:g} keV, outside [1, 150] keV")
    rows = np.zeros((len(spectra), len(grid)))
    for k, (e, w) in enumerate(spectra):
        e, w = np.asarray(e, dtype=float), np.asarray(w, dtype=float)
        on = w > 0
        np.add.at(rows[k], np.searchsorted(grid, e[on]), w[on])
    for k in range(len(spectra)):
        if rel[k] <= 0 or rows[k].sum() <= 0:
            rows[k] = rows[nz[np.argmin(np.abs(nz - k))]]
    cdf = np.cumsum(rows, axis=1)
    cdf /= cdf[:, -1:]
    return grid, cdf, rel, int(nz[0]), int(nz[-1])


def photon_count_through_field(src: dict) -> float:
    """Photons crossing the field for source.mas (chatcarlo/source.py:213-258): the fluence per
    mAs at the SID times mAs times the field area.  Without the heel effect the fluence is the
    central-axis value (spectrum.fluence_per_mas); with it, the mean over a 101 x 101 grid of the
    field (inside the circle for a cone) of the heel bins' fluence per mAs interpolated at each
    point's heel coordinate.  SpekPy supplies both and is absent here, so this raises, as the
    reference does without SpekPy."""
    mas = src.get("mas")
    if mas is None:
        raise ValueError("source.mas is not given: the photon-count calibration needs the tube current-time "
                         "product in mAs")
    fld = src["field"]
    sid = float(fld["sid_cm"])
    cone = fld.get("shape", "rect") == "cone"
    area = math.pi * (fld["diameter_cm"] / 2.0) ** 2 if cone else float(fld["size_cm"][0]) * float(fld["size_cm"][1])
    if not src.get("heel_effect"):
        return _spectrum.fluence_per_mas(src, sid) * float(mas) * area
    d = np.asarray(src["direction"], dtype=float)
    d = d / np.linalg.norm(d)
    u, v = beam_basis(d)
    hu, hv = heel_axis_coeffs(src, d, u, v)
    span = heel_span_cm(fld, hu, hv)
    if cone:
        g = np.linspace(-span, span, 101)
        gu, gv = np.meshgrid(g, g)
        s = (gu * hu + gv * hv)[gu ** 2 + gv ** 2 <= span ** 2]
    else:
        w, h = (float(x) for x in fld["size_cm"])
        gu, gv = np.meshgrid(np.linspace(-w / 2, w / 2, 101), np.linspace(-h / 2, h / 2, 101))
        s = (gu * hu + gv * hv).ravel()
    centres, _, _, flu_abs = _spectrum.heel_spectra_for_source(src, sid, span)
    return float(np.interp(s, centres, flu_abs).mean()) * float(mas) * area


class SourceRecord:
    """A cc_source plus the arrays it points to (kept alive with it)."""

    def __init__(self, src: dict):
        d = np.asarray(src["direction"], dtype=float)
        d = d / np.linalg.norm(d)
        u, v = beam_basis(d)
        fld = src["field"]
        shape = fld.get("shape", "rect")
        heel = None
        if src.get("heel_effect"):
            if shape == "parallel":
                raise ValueError("a parallel field cannot take source.heel_effect")
            hu, hv = heel_axis_coeffs(src, d, u, v)
            span = heel_span_cm(fld, hu, hv)
            centres, spectra, rel, _ = _spectrum.heel_spectra_for_source(src, float(fld["sid_cm"]), span)
            grid, cdf, rel, lo, hi = heel_tables(centres, spectra, rel)
            heel = (hu, hv, span, rel, cdf, lo, hi)
            # the plain spectrum fields: the bin nearest the central axis (unused by the sampler)
            mid = min(max(len(rel) // 2, lo), hi)
            self.spec_e, self.spec_cdf = grid, cdf[mid]
        else:
            self.spec_e, self.spec_cdf = spectrum_table(src)
        self.spec_e = np.ascontiguousarray(self.spec_e, dtype=np.float64)
        self.spec_cdf = np.ascontiguousarray(self.spec_cdf, dtype=np.float64)
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
        if heel is not None:
            hu, hv, span, rel, cdf, lo, hi = heel
            self.heel_rel_flu = np.ascontiguousarray(rel, dtype=np.float64)
            self.heel_cdf = np.ascontiguousarray(cdf, dtype=np.float64)
            s.heel_bins, s.heel_nz_lo, s.heel_nz_hi = len(rel), lo, hi
            s.heel_hu, s.heel_hv, s.heel_span = hu, hv, span
            s.heel_rel_flu = N.ptr(self.heel_rel_flu, N.C.c_double)
            s.heel_cdf = N.ptr(self.heel_cdf, N.C.c_double)
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
