"""Geometry.nearest_object_distance_cm, the host-side distance the diagnostics' near-source warning
uses: the reference's cases (tests/test_geometry.py:9-41), and the world box around the scene."""
import numpy as np

from chatcarlo_amd.geometry import Geometry

BOX = {"name": "box", "shape": "box", "material": "water", "center": [0.0, 0.0, 0.0], "size_cm": [10.0, 10.0, 10.0]}


def test_nearest_object_distance_outside_box():
    assert np.isclose(Geometry([BOX]).nearest_object_distance_cm([15.0, 0.0, 0.0]), 10.0)


def test_nearest_object_distance_zero_when_inside():
    assert Geometry([BOX]).nearest_object_distance_cm([0.0, 0.0, 0.0]) == 0.0


def test_nearest_object_distance_picks_closest_of_multiple():
    geom = Geometry([
        {"name": "near", "shape": "sphere", "material": "water", "center": [10.0, 0.0, 0.0], "radius_cm": 2.0},
        {"name": "far", "shape": "sphere", "material": "water", "center": [-100.0, 0.0, 0.0], "radius_cm": 2.0},
    ])
    assert np.isclose(geom.nearest_object_distance_cm([0.0, 0.0, 0.0]), 8.0)


def test_nearest_object_distance_none_when_no_objects():
    geom = Geometry.__new__(Geometry)
    geom.geoms = []
    geom.background = "air"
    assert geom.nearest_object_distance_cm([0.0, 0.0, 0.0]) is None


def test_world_box_is_the_scene_bbox_with_the_margin():
    """chatcarlo/geometry.py:160-164: the bounding box of every primitive, 50 cm either side."""
    geom = Geometry([BOX, {"name": "rod", "shape": "cylinder", "material": "bone", "center": [20.0, 0.0, 0.0],
                           "radius_cm": 1.5, "height_cm": 40.0, "axis": "z"}])
    np.testing.assert_allclose(geom.bbox_min, [-55.0, -55.0, -70.0])
    np.testing.assert_allclose(geom.bbox_max, [71.5, 55.0, 70.0])
