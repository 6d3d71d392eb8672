"""CTDI calibration and dose diagnostics, host side (chatcarlo/ctdi.py:53-131,
chatcarlo/diagnostics.py:16-67): the CTDIw weighting and hole selection on a synthetic dose
field, the phantom geometry, and the warning rules."""
import numpy as np

from chatcarlo_amd.ctdi import PHANTOM_DIAMETER_CM, ctdi_from_dose, ctdi_phantom_geometry
from chatcarlo_amd.diagnostics import background_medium_warning, max_voxel_position_cm, near_source_air_warning
from chatcarlo_amd.tally import VoxelGrid


def test_ctdiw_weights_and_hole_selection():
    # body phantom about the y axis at (0, 0, 110): centre hole r < 2, ring 14 <= r <= 15, |y| <= 5
    iso = np.array([0.0, 0.0, 110.0])
    grid = VoxelGrid.from_bbox(iso - 18.0, iso + 18.0, 1.0)
    c = grid.voxel_centers().reshape(*grid.shape, 3)
    r = np.hypot(c[..., 0] - iso[0], c[..., 2] - iso[2])
    dose = np.full(grid.shape, 100.0)
    inside = np.abs(c[..., 1]) <= 5.0
    dose[inside & (r < 2.0)] = 1.0
    dose[inside & (r >= 14.0) & (r <= 15.0)] = 4.0
    dose[~inside] = 1e6          # outside the chamber length: must not count
    ctdiw, dc, dp = ctdi_from_dose(dose, grid, iso, 1, 16.0)
    assert dc == 1.0 and dp == 4.0
    assert abs(ctdiw - 3.0) < 1e-12


def test_phantom_geometry():
    g = ctdi_phantom_geometry({"isocenter": [1, 2, 3], "axis": "x"}, "head")
    assert g.geoms[0]["radius_cm"] == PHANTOM_DIAMETER_CM["head"] / 2 == 8.0
    assert g.geoms[0]["axis"] == "x" and g.geoms[0]["material"] == "pmma" and g.geoms[0]["height_cm"] == 15.0
    assert g.geoms[0]["center"] == [1.0, 2.0, 3.0]


def test_warning_rules():
    assert background_medium_warning("water", "air") is None
    assert "air" in background_medium_warning("air", "air")
    assert near_source_air_warning("water", "air", 1.0, 50.0) is None
    assert near_source_air_warning("air", "air", 60.0, 50.0) is None
    assert near_source_air_warning("air", "air", 1.0, None) is None
    assert "1/r^2" in near_source_air_warning("air", "air", 10.0, 50.0)


def test_max_voxel_position():
    grid = VoxelGrid(origin_cm=np.array([-1.0, 0.0, 2.0]), shape=(3, 4, 5), voxel_size_cm=2.0)
    data = np.zeros(grid.shape)
    data[2, 1, 4] = 7.0
    np.testing.assert_allclose(max_voxel_position_cm(grid, data), [-1 + 5.0, 3.0, 2 + 9.0])
