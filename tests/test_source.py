"""Source geometry: the closed-form checks of the reference's tests/test_cone_beam.py,
tests/test_ct_rotation.py and tests/test_parallel_beam.py, applied to the oracle's sampler (CPU)
and to the device sampler (`cc_test_source`, marked gpu), plus the scene-validation cases of those
files.  The energy spectrum is one line: these tests look at positions and directions only."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle.source import sample_source

CONE = {"kvp": 100.0, "filtration_mm_al": 2.5, "position": [0.0, 0.0, 170.0], "direction": [0.0, 0.0, -1.0],
        "field": {"shape": "cone", "diameter_cm": 40.0, "sid_cm": 60.0}}
ISO = np.array([0.0, 0.0, 110.0])
CT = {"kvp": 120.0, "filtration_mm_al": 2.5, "position": [0.0, 0.0, 170.0], "direction": [0.0, 0.0, -1.0],
      "rotation": {"isocenter": ISO.tolist(), "axis": "y", "n_angles": 36},
      "field": {"size_cm": [50.0, 4.0], "sid_cm": 60.0}}
PAR = {"spectrum": [{"energy_keV": 60.0, "weight": 1.0}], "position": [0.0, 0.0, 15.0],
       "direction": [0.0, 1.0, 0.0], "field": {"shape": "parallel", "size_cm": [10.0, 10.0]}}
LINE = (np.array([60.0]), np.array([1.0]))


def oracle_sampler(src, n, seed):
    return sample_source(src, n, np.random.default_rng(seed), *LINE)[:2]


def device_sampler(src, n, seed):
    from chatcarlo_amd import _native as N
    from chatcarlo_amd.source import SourceRecord
    from chatcarlo_amd.transport import default_engine
    eng = default_engine()
    rec = SourceRecord({**src, "spectrum": [{"energy_keV": 60.0, "weight": 1.0}]})
    pos = np.zeros((n, 3), np.float32); dr = np.zeros((n, 3), np.float32); en = np.zeros(n, np.float32)
    N.check(eng.lib.cc_test_source(eng.ctx, C.byref(rec.struct), n, seed, 0, N.ptr(pos, C.c_float),
                                   N.ptr(dr, C.c_float), N.ptr(en, C.c_float)))
    assert np.all(en == 60.0)
    return pos.astype(np.float64), dr.astype(np.float64)


SAMPLERS = [pytest.param(oracle_sampler, id="oracle"),
            pytest.param(device_sampler, id="device", marks=pytest.mark.gpu)]


def _half_cos(fld):
    return np.cos(np.arctan(fld["diameter_cm"] / 2.0 / fld["sid_cm"]))


@pytest.mark.parametrize("sampler", SAMPLERS)
def test_cone_directions_unit_and_within_the_half_angle(sampler):
    """tests/test_cone_beam.py:19-25."""
    _, d = sampler(CONE, 20_000, 41)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    assert np.all(d @ np.array([0.0, 0.0, -1.0]) >= _half_cos(CONE["field"]) - 1e-6)


@pytest.mark.parametrize("sampler", SAMPLERS)
def test_cone_directions_uniform_in_solid_angle(sampler):
    """cos theta uniform on [cos half-angle, 1], azimuth uniform (tests/test_cone_beam.py:28-41)."""
    n = 100_000
    _, d = sampler(CONE, n, 43)
    ch = _half_cos(CONE["field"])
    c = d @ np.array([0.0, 0.0, -1.0])
    width = 1.0 - ch
    se = width / np.sqrt(12.0) / np.sqrt(n)
    assert abs(c.mean() - (1.0 + ch) / 2.0) < 5 * se
    assert abs(c.std() - width / np.sqrt(12.0)) < 5 * se
    assert abs(d[:, 0].mean()) < 5e-3 and abs(d[:, 1].mean()) < 5e-3


@pytest.mark.parametrize("sampler", SAMPLERS)
def test_cone_composes_with_rotation(sampler):
    """Every direction within the half-angle of its rotated beam axis, focus to isocentre
    (tests/test_cone_beam.py:44-56)."""
    src = {**CONE, "rotation": {"isocenter": ISO.tolist(), "axis": "y"}}
    p, d = sampler(src, 20_000, 47)
    ax = ISO[None, :] - p
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    assert np.all(np.sum(d * ax, axis=1) >= _half_cos(src["field"]) - 1e-6)


@pytest.mark.parametrize("sampler", SAMPLERS)
def test_rotating_source_on_the_isocentre_circle_at_discrete_angles(sampler):
    """Focus on the 60 cm circle about the y axis through the isocentre, all 36 angles used, beam
    pointing at the isocentre side (tests/test_ct_rotation.py:23-54)."""
    p, d = sampler(CT, 20_000, 11)
    assert np.allclose(p[:, 1], ISO[1], atol=1e-4)
    assert np.allclose(np.linalg.norm(p[:, [0, 2]] - ISO[[0, 2]], axis=1), 60.0, atol=1e-4)
    theta = np.arctan2(p[:, 0] - ISO[0], p[:, 2] - ISO[2])
    k = np.round(theta / (2 * np.pi / 36)).astype(int) % 36
    assert len(set(k.tolist())) == 36
    assert np.allclose(theta, np.round(theta / (2 * np.pi / 36)) * (2 * np.pi / 36), atol=1e-5)
    to_iso = ISO[None, :] - p
    to_iso /= np.linalg.norm(to_iso, axis=1, keepdims=True)
    assert np.all(np.sum(d * to_iso, axis=1) > 0.9)


@pytest.mark.parametrize("sampler", SAMPLERS)
def test_no_rotation_is_a_static_focus(sampler):
    """tests/test_ct_rotation.py:57-61."""
    src = {k: v for k, v in CT.items() if k != "rotation"}
    p, _ = sampler(src, 100, 17)
    assert np.allclose(p, np.asarray(src["position"]), atol=1e-5)


@pytest.mark.parametrize("sampler", SAMPLERS)
def test_continuous_rotation_without_n_angles(sampler):
    """Every one of 720 angle bins filled, radius unchanged (tests/test_ct_rotation.py:64-75)."""
    src = {**CT, "rotation": {"isocenter": ISO.tolist(), "axis": "y"}}
    p, _ = sampler(src, 50_000, 19)
    theta = np.arctan2(p[:, 0] - ISO[0], p[:, 2] - ISO[2])
    hist, _ = np.histogram(theta, bins=720, range=(-np.pi, np.pi))
    assert np.all(hist > 0)
    assert np.allclose(np.linalg.norm(p[:, [0, 2]] - ISO[[0, 2]], axis=1), 60.0, atol=1e-4)


@pytest.mark.parametrize("sampler", SAMPLERS)
def test_helical_scan_translates_uniformly_along_the_axis(sampler):
    """Shift along the rotation axis uniform on [-L/2, L/2] (tests/test_ct_rotation.py:78-97)."""
    scan, n = 30.0, 50_000
    src = {**CT, "rotation": {"isocenter": ISO.tolist(), "axis": "y", "scan_length_cm": scan}}
    p, _ = sampler(src, n, 23)
    y = p[:, 1] - ISO[1]
    assert y.min() >= -scan / 2 - 1e-5 and y.max() <= scan / 2 + 1e-5
    assert abs(y.mean()) < 5 * scan / np.sqrt(12.0) / np.sqrt(n)
    assert abs(y.std() - scan / np.sqrt(12.0)) < 0.2
    assert np.allclose(np.linalg.norm(p[:, [0, 2]] - ISO[[0, 2]], axis=1), 60.0, atol=1e-4)


@pytest.mark.parametrize("sampler", SAMPLERS)
def test_parallel_beam(sampler):
    """One direction for all photons; origins uniform over size_cm on the focus plane
    (tests/test_parallel_beam.py:18-36)."""
    n = 100_000
    p, d = sampler(PAR, n, 2)
    assert np.allclose(d, [0.0, 1.0, 0.0], atol=1e-7)
    assert np.allclose(p[:, 1], 0.0, atol=1e-6)
    assert p[:, 0].min() >= -5.0 - 1e-5 and p[:, 0].max() <= 5.0 + 1e-5
    assert p[:, 2].min() >= 10.0 - 1e-5 and p[:, 2].max() <= 20.0 + 1e-5
    se = 10.0 / np.sqrt(12.0) / np.sqrt(n)
    assert abs(p[:, 0].mean()) < 5 * se and abs(p[:, 2].mean() - 15.0) < 5 * se


def _scene(field):
    return {"source": {"kvp": 100, "position": [0, 0, 170], "direction": [0, 0, -1], "field": field},
            "geometry": [{"name": "p", "shape": "sphere", "material": "water", "radius_cm": 10,
                          "center": [0, 0, 110]}]}


def test_scene_validation_of_field_shapes():
    """tests/test_cone_beam.py:78-94 and tests/test_parallel_beam.py:48-55: a cone needs its
    diameter, unknown shapes are rejected, a parallel beam needs no SID."""
    from chatcarlo_amd.scene import validate_scene
    assert not validate_scene(_scene({"shape": "cone", "sid_cm": 60})).ok
    assert not validate_scene(_scene({"shape": "hexagon", "size_cm": [10, 10], "sid_cm": 60})).ok
    assert validate_scene(_scene({"shape": "cone", "diameter_cm": 40, "sid_cm": 60})).ok
    sc = validate_scene({"source": {**PAR}, "geometry": [{"name": "p", "shape": "box", "material": "water",
                                                         "size_cm": [30, 20, 30], "center": [0, 10, 15]}]})
    assert sc.ok, sc.errors
