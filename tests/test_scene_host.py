"""Scene validation of source spectra, kVp and field shapes: the cases of the reference's
tests/test_scene.py:24-118, same inputs, same verdicts and error paths."""
from __future__ import annotations

import pytest

from chatcarlo_amd.scene import validate_scene

BASE = {"geometry": [{"name": "phantom", "shape": "box", "material": "water", "size_cm": [30, 30, 20],
                      "center": [0, 0, 10]}]}
LINE60 = [{"energy_keV": 60.0, "weight": 1.0}]


def _src(**over):
    src = {"position": [0, -100, 10], "direction": [0, 1, 0],
           "field": {"shape": "rect", "size_cm": [10, 10], "sid_cm": 100}}
    src.update(over)
    return {**BASE, "source": src}


@pytest.mark.parametrize("over,ok,where", [
    ({}, False, ("path", "kvp")),
    ({"spectrum": LINE60}, True, None),
    ({"spectrum": [{"energy_keV": -1.0, "weight": 1.0}]}, False, ("path=", "source.spectrum")),
    ({"spectrum": [{"energy_keV": 200.0, "weight": 1.0}]}, False, ("path=", "source.spectrum")),
    ({"spectrum": [{"energy_keV": 150.0, "weight": 1.0}]}, True, None),
    ({"kvp": 180.0}, False, ("path", "kvp")),
    ({"spectrum": LINE60, "mas": 4.0}, False, ("message", "mas")),
    ({"spectrum": LINE60, "heel_effect": True, "anode_direction": [1, 0, 0]}, False, ("message", "heel_effect")),
    ({"spectrum": LINE60, "ctdi_vol_mGy": 10.0, "rotation": {"isocenter": [0, 0, 10]}}, False,
     ("message", "ctdi_vol_mGy")),
    ({"kvp": 80.0, "spectrum": LINE60}, False, ("message", "kvp")),
    ({"spectrum": LINE60, "field": {"shape": "parallel", "size_cm": [10, 10]}}, True, None),
    ({"spectrum": LINE60, "field": {"shape": "parallel"}}, False, ("path", "size_cm")),
    ({"kvp": 100.0, "mas": 4.0, "field": {"shape": "parallel", "size_cm": [10, 10]}}, False, ("message", "mas")),
    ({"kvp": 100.0, "heel_effect": True, "anode_direction": [1, 0, 0],
      "field": {"shape": "parallel", "size_cm": [10, 10]}}, False, ("message", "heel_effect")),
])
def test_source_validation_matches_the_reference(over, ok, where):
    sc = validate_scene(_src(**over))
    assert sc.ok == ok, sc.errors
    if where is not None:
        kind, text = where
        if kind == "path=":
            assert any(e.path == text for e in sc.errors), sc.errors
        elif kind == "path":
            assert any(text in e.path for e in sc.errors), sc.errors
        else:
            assert any(text in e.message for e in sc.errors), sc.errors


def test_parallel_field_refuses_any_given_mas():
    """The reference tests `source.mas is not None` for a parallel field (chatcarlo/scene.py:140-147):
    mas: 0 is refused for the combination, besides failing its own positivity check."""
    sc = validate_scene(_src(kvp=100.0, mas=0, field={"shape": "parallel", "size_cm": [10, 10]}))
    assert any(e.path == "source" and "parallel" in e.message and "mas" in e.message for e in sc.errors)
    assert any(e.path == "source.mas" for e in sc.errors)


def test_explicit_nulls_are_checked_and_defaults_fill_absent_keys():
    """dict.get(key, default) semantics of the reference: a default fills an absent key only."""
    sc = validate_scene(_src(kvp=80.0, filtration_mm_al=None))
    assert any(e.path == "source.filtration_mm_al" for e in sc.errors)
    sc = validate_scene(_src(kvp=80.0, anode_angle_deg=None))
    assert sc.ok and any(w.path == "source.anode_angle_deg" for w in sc.warnings)
    assert sc.raw["source"]["anode_angle_deg"] == 12.0
    sc = validate_scene(_src(kvp=80.0, rotation={"isocenter": [0, 0, 10], "axis": None}))
    assert any(e.path == "source.rotation.axis" for e in sc.errors)


def test_normalised_values_are_written_back():
    raw = _src(kvp=80, anode_angle_deg=10, ctdi_vol_mGy=12, rotation={"isocenter": [0, 0, 10], "scan_length_cm": 8})
    raw["geometry"].append({"name": "rod", "shape": "cylinder", "material": "bone", "center": [0, 0, 10],
                            "radius_cm": 1, "height_cm": 10})
    sc = validate_scene(raw)
    assert sc.ok, sc.errors
    s = sc.raw["source"]
    assert s["anode_angle_deg"] == 10.0 and isinstance(s["anode_angle_deg"], float)
    assert s["ctdi_vol_mGy"] == 12.0 and isinstance(s["ctdi_vol_mGy"], float) and s["ctdi_phantom"] == "body"
    assert s["rotation"]["axis"] == "z" and isinstance(s["rotation"]["scan_length_cm"], float)
    assert s["field"]["shape"] == "rect"
    assert sc.raw["geometry"][1]["axis"] == "z"
    sc = validate_scene({**raw, "physics": {}})
    assert sc.raw["physics"]["fluorescence"] is True
    sc = validate_scene({**raw, "physics": {"fluorescence": False, "pairs": 1}})
    assert any(e.path == "physics" and "pairs" in e.message for e in sc.errors)


def test_geometry_rules():
    raw = _src(kvp=80.0)
    raw["geometry"] = [{"shape": "box", "material": "water", "center": [0, 0, 0], "size_cm": [1, 1, 1]},
                       {"name": "object_0", "shape": "sphere", "material": "", "center": [0, 0], "radius_cm": 0},
                       "not a mapping"]
    sc = validate_scene(raw)
    paths = {e.path for e in sc.errors}
    assert {"geometry[1].name", "geometry[1].material", "geometry[1].center", "geometry[1].radius_cm",
            "geometry[2]"} <= paths
    assert not validate_scene({**raw, "geometry": []}).ok
    sc = validate_scene(_src(kvp=80.0, position=[0, 0, 10]))
    assert sc.ok and any(w.path == "source.position" for w in sc.warnings)
