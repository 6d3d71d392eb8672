"""Dose-grid post-processing and the checks on where the maxima fall
(drop-in for chatcarlo/diagnostics.py:16-67).

The maximum absorbed dose or H*(10) of a `--dose-grid` run can sit in a background (air)
voxel, where kerma / density is inflated by the low density, or next to the point source,
where the 1/r^2 fluence of the point-source model diverges.  The run command reports both
cases with these checks.
"""
from __future__ import annotations

import numpy as np

from .geometry import Geometry
from .materials import density
from .tally import VoxelGrid


def dose_map_Gy(grid: VoxelGrid, geometry: Geometry) -> np.ndarray:
    """Kerma -> absorbed dose [Gy] with the density of the material at each voxel centre."""
    mats = geometry.material_at(grid.voxel_centers())
    dens = {m: density(m) for m in set(mats.tolist())}
    rho = np.array([dens[m] for m in mats], dtype=float).reshape(grid.shape)
    return grid.dose_map_Gy(rho)


def max_voxel_position_cm(grid: VoxelGrid, data: np.ndarray) -> np.ndarray:
    """Centre [cm] of the voxel holding the maximum of `data`."""
    idx = np.unravel_index(int(np.argmax(data)), data.shape)
    return grid.origin_cm + (np.asarray(idx, dtype=float) + 0.5) * grid.voxel_size_cm


def background_medium_warning(material: str, background: str) -> str | None:
    """A warning when the maximum-dose voxel lies in the background medium."""
    if material != background:
        return None
    return ("the maximum lies in a background-medium (air) voxel; absorbed dose is specific to the "
            "medium, so this value says nothing about the exposure of a patient or detector")


def near_source_air_warning(material: str, background: str, distance_from_source_cm: float,
                            nearest_object_distance_cm: float | None) -> str | None:
    """A warning when the H*(10) maximum is a background voxel closer to the source than any
    object of the scene: the 1/r^2 divergence of the point-source model, not a real position."""
    if material != background:
        return None
    if nearest_object_distance_cm is None or distance_from_source_cm >= nearest_object_distance_cm:
        return None
    return (f"the maximum lies in an air voxel {distance_from_source_cm:.1f} cm from the source, closer "
            f"than any object of the scene (nearest {nearest_object_distance_cm:.1f} cm): the 1/r^2 "
            "divergence of the point source, not a physical exposure; score the position of interest "
            "with a fine grid of its own")
