"""Oracle source sampling (chatcarlo/source.py:20-35, 67-210).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
Pinned by the closed-form checks of the reference's tests/test_cone_beam.py,
test_ct_rotation.py and test_parallel_beam.py (cone opening, rotation radius,
helical shift range, parallel-beam footprint), restated in tests/test_source.py.
The heel effect (sample_source_heel) is pinned in tests/test_source.py by the
closed forms a rect field gives: the kept heel coordinates follow the
interpolated relative fluence, and each heel bin's energies its own spectrum.
"""
from __future__ import annotations

import math

import numpy as np

AXES = {"x": np.array([1.0, 0.0, 0.0]), "y": np.array([0.0, 1.0, 0.0]), "z": np.array([0.0, 0.0, 1.0])}


def rotate_batch(v, axis, angles):
    """Rodrigues rotation of a fixed vector, chatcarlo/source.py:20-25."""
    ca = np.cos(angles)[:, None]
    sa = np.sin(angles)[:, None]
    return v[None, :] * ca + np.cross(axis, v)[None, :] * sa + axis[None, :] * (axis @ v) * (1.0 - ca)


def beam_basis(d):
    """chatcarlo/source.py:28-35."""
    if abs(d[2]) < 0.999:
        u = np.array([-d[1], d[0], 0.0])
    else:
        u = np.array([1.0, 0.0, 0.0])
    u = u / np.linalg.norm(u)
    return u, np.cross(d, u)


def sample_source(src, n, rng, spec_e, spec_cdf):
    """Positions, directions, energies; uniform order as chatcarlo/source.py:104-209."""
    pos = np.asarray(src["position"], dtype=float)
    d = np.asarray(src["direction"], dtype=float)
    d = d / np.linalg.norm(d)
    fld = src["field"]
    sid = fld.get("sid_cm")
    u, v = beam_basis(d)
    rot = src.get("rotation")
    if rot is not None:
        iso = np.asarray(rot["isocenter"], dtype=float)
        axis = AXES[rot.get("axis", "z")]
        na = rot.get("n_angles")
        if na:
            angles = 2.0 * np.pi * rng.integers(0, int(na), size=n) / int(na)
        else:
            angles = rng.uniform(0.0, 2.0 * np.pi, size=n)
        pos_a = iso[None, :] + rotate_batch(pos - iso, axis, angles)
        d_a, u_a, v_a = (rotate_batch(x, axis, angles) for x in (d, u, v))
        scan = float(rot.get("scan_length_cm") or 0.0)
        if scan > 0.0:
            shift = rng.uniform(-scan / 2.0, scan / 2.0, size=n)
            pos_a = pos_a + axis[None, :] * shift[:, None]
        origins = pos_a
    else:
        pos_a, d_a, u_a, v_a = pos, d, u, v
        origins = np.tile(pos, (n, 1))
    shape = fld.get("shape", "rect")
    if shape == "cone":
        cos_half = math.cos(math.atan((fld["diameter_cm"] / 2.0) / sid))
        c = rng.uniform(cos_half, 1.0, n)
        ph = rng.uniform(0.0, 2.0 * np.pi, n)
        s = np.sqrt(1.0 - c * c)
        dirs = c[:, None] * d_a + (s * np.cos(ph))[:, None] * u_a + (s * np.sin(ph))[:, None] * v_a
    elif shape == "parallel":
        w, h = fld["size_cm"]
        su = rng.uniform(-w / 2, w / 2, n)
        sv = rng.uniform(-h / 2, h / 2, n)
        origins = origins + su[:, None] * u_a + sv[:, None] * v_a
        dirs = np.broadcast_to(d_a, (n, 3)).copy()
    else:
        w, h = fld["size_cm"]
        a = rng.uniform(-w / 2, w / 2, n)
        b = rng.uniform(-h / 2, h / 2, n)
        target = pos_a + d_a * sid + a[:, None] * u_a + b[:, None] * v_a
        dirs = target - origins
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    # rng.choice(p) == searchsorted(cdf, U, side="right") (chatcarlo/spectrum.py:82)
    k = np.searchsorted(spec_cdf, rng.random(n), side="right")
    energies = spec_e[np.minimum(k, len(spec_e) - 1)]
    return np.array(origins, dtype=float), dirs, energies


def source_from_uniforms(src, u, spec_e, spec_cdf):
    """The GPU source sampler's slot order restated in FP64 (one row of uniforms per history).

    Slots: u0 angle, u1 helical shift, u2 field a, u3 field b, u4 energy, u5 tau.
    Same formulas as sample_source above (chatcarlo/source.py:104-209).
    """
    u = np.asarray(u, dtype=float)
    n = u.shape[0]
    pos = np.asarray(src["position"], dtype=float)
    d = np.asarray(src["direction"], dtype=float)
    d = d / np.linalg.norm(d)
    fld = src["field"]
    sid = fld.get("sid_cm")
    uu, vv = beam_basis(d)
    rot = src.get("rotation")
    if rot is not None:
        iso = np.asarray(rot["isocenter"], dtype=float)
        axis = AXES[rot.get("axis", "z")]
        na = rot.get("n_angles")
        if na:
            k = np.minimum((u[:, 0] * na).astype(np.int64), na - 1)
            angles = 2.0 * np.pi * k / na
        else:
            angles = 2.0 * np.pi * u[:, 0]
        pos_a = iso[None, :] + rotate_batch(pos - iso, axis, angles)
        d_a, u_a, v_a = (rotate_batch(x, axis, angles) for x in (d, uu, vv))
        scan = float(rot.get("scan_length_cm") or 0.0)
        if scan > 0.0:
            pos_a = pos_a + axis[None, :] * (-scan / 2.0 + scan * u[:, 1])[:, None]
    else:
        pos_a = np.tile(pos, (n, 1))
        d_a, u_a, v_a = (np.tile(x, (n, 1)) for x in (d, uu, vv))
    shape = fld.get("shape", "rect")
    if shape == "cone":
        cos_half = math.cos(math.atan((fld["diameter_cm"] / 2.0) / sid))
        c = cos_half + (1.0 - cos_half) * u[:, 2]
        ph = 2.0 * np.pi * u[:, 3]
        s = np.sqrt(1.0 - c * c)
        dirs = c[:, None] * d_a + (s * np.cos(ph))[:, None] * u_a + (s * np.sin(ph))[:, None] * v_a
        origins = pos_a
    elif shape == "parallel":
        w, h = fld["size_cm"]
        a = -w / 2 + w * u[:, 2]
        b = -h / 2 + h * u[:, 3]
        origins = pos_a + a[:, None] * u_a + b[:, None] * v_a
        dirs = d_a.copy()
    else:
        w, h = fld["size_cm"]
        a = -w / 2 + w * u[:, 2]
        b = -h / 2 + h * u[:, 3]
        t = d_a * sid + a[:, None] * u_a + b[:, None] * v_a
        dirs = t / np.linalg.norm(t, axis=1, keepdims=True)
        origins = pos_a
    k = np.searchsorted(spec_cdf, u[:, 4], side="right")
    energies = spec_e[np.minimum(k, len(spec_e) - 1)]
    return origins, dirs, energies


def sample_source_heel(src, n, rng, hu, hv, span, centres, spectra, rel_flu):
    """Source photons with the heel effect, as chatcarlo/source.py:128-149,195-207 draws them:
    field points (rect: offsets a, b; cone: cos theta, phi) are drawn in rounds and each kept
    with probability interp(s, centres, rel_flu) at its heel coordinate s, until n are kept;
    each photon's energy comes from the spectrum of its bin of [-span, span] (clamped to the
    bins of non-zero fluence).  Returns origins, directions, energies and the heel coordinates."""
    pos = np.asarray(src["position"], dtype=float)
    d = np.asarray(src["direction"], dtype=float)
    d = d / np.linalg.norm(d)
    fld = src["field"]
    sid = float(fld["sid_cm"])
    u, v = beam_basis(d)
    rot = src.get("rotation")
    if rot is not None:
        iso = np.asarray(rot["isocenter"], dtype=float)
        axis = AXES[rot.get("axis", "z")]
        angles = rng.uniform(0.0, 2.0 * np.pi, size=n)
        origins = iso[None, :] + rotate_batch(pos - iso, axis, angles)
        d_a, u_a, v_a = (rotate_batch(x, axis, angles) for x in (d, u, v))
    else:
        origins = np.tile(pos, (n, 1))
        d_a, u_a, v_a = (np.tile(x, (n, 1)) for x in (d, u, v))
    cone = fld.get("shape", "rect") == "cone"
    if cone:
        cos_half = math.cos(math.atan((fld["diameter_cm"] / 2.0) / sid))
    else:
        w, h = fld["size_cm"]
    kept = [np.empty(0)] * 3
    while len(kept[0]) < n:
        m = max(int((n - len(kept[0])) * 1.8), 1024)
        if cone:
            p1 = rng.uniform(cos_half, 1.0, m)
            p2 = rng.uniform(0.0, 2.0 * np.pi, m)
            s = sid * np.sqrt(1.0 - p1 * p1) / p1 * (np.cos(p2) * hu + np.sin(p2) * hv)
        else:
            p1 = rng.uniform(-w / 2, w / 2, m)
            p2 = rng.uniform(-h / 2, h / 2, m)
            s = p1 * hu + p2 * hv
        keep = rng.random(m) < np.interp(s, centres, rel_flu)
        kept = [np.concatenate([a, b[keep]]) for a, b in zip(kept, (s, p1, p2))]
    s, p1, p2 = (a[:n] for a in kept)
    if cone:
        sn = np.sqrt(1.0 - p1 * p1)
        dirs = p1[:, None] * d_a + (sn * np.cos(p2))[:, None] * u_a + (sn * np.sin(p2))[:, None] * v_a
    else:
        t = d_a * sid + p1[:, None] * u_a + p2[:, None] * v_a
        dirs = t / np.linalg.norm(t, axis=1, keepdims=True)
    nb = len(centres)
    bins = np.clip(((s + span) / (2.0 * span / nb)).astype(int), 0, nb - 1)
    nz = np.flatnonzero(np.asarray(rel_flu) > 0)
    bins = np.clip(bins, nz.min(), nz.max())
    energies = np.empty(n)
    for b in np.unique(bins):
        sel = bins == b
        e_b, w_b = spectra[b]
        energies[sel] = np.asarray(e_b)[rng.choice(len(e_b), size=int(sel.sum()), p=np.asarray(w_b))]
    return origins, dirs, energies, s
