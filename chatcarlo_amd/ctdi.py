"""Absolute calibration from the console CTDIvol (drop-in for chatcarlo/ctdi.py:43-131).

The scene's source (kVp, filtration, rotation, field) irradiates a standard CTDI phantom (PMMA
cylinder, body 32 cm or head 16 cm diameter, 15 cm long, at the rotation centre) on the GPU; the
per-history weighted CTDI

    CTDIw = D_center / 3 + 2 D_periphery / 3

uses the mean dose of the voxels of the central hole (within 2 cm of the axis) and of the
peripheral ring (1-2 cm below the surface), both over the 100 mm chamber length.  The measured
CTDIvol divided by it is the effective photon count that scales per-history results to absolute
ones (TransportResult.n_photons_real).  The scene's own geometry is not used: the scatter of a
table or walls is not part of a measured CTDI either.
"""
from __future__ import annotations

import numpy as np

from .geometry import Geometry
from .tally import VoxelGrid

PHANTOM_DIAMETER_CM = {"body": 32.0, "head": 16.0}
PHANTOM_LENGTH_CM = 15.0
CHAMBER_LENGTH_CM = 10.0
GRID_RESOLUTION_CM = 1.0
_AXIS = {"x": 0, "y": 1, "z": 2}


def ctdi_phantom_geometry(rot: dict, phantom: str = "body") -> Geometry:
    """The PMMA CTDI phantom centred on the rotation centre, its axis the rotation axis."""
    return Geometry([{"name": f"ctdi_phantom_{phantom}", "shape": "cylinder", "material": "pmma",
                      "axis": rot.get("axis", "z"), "radius_cm": PHANTOM_DIAMETER_CM[phantom] / 2.0,
                      "height_cm": PHANTOM_LENGTH_CM, "center": [float(x) for x in rot["isocenter"]]}])


def ctdi_per_history_Gy(src: dict, phantom: str = "body", n_histories: int = 200_000,
                        seed: int | None = None, engine=None) -> tuple[float, float, float]:
    """(CTDIw, D_center, D_periphery) in Gy per history."""
    from . import _native as N
    from .diagnostics import dose_map_Gy
    from .source import SourceRecord
    from .transport import _seed_from, default_engine
    rot = src.get("rotation")
    if rot is None:
        raise ValueError("CTDIvol calibration needs source.rotation (a gantry rotation)")
    geometry = ctdi_phantom_geometry(rot, phantom)
    radius = PHANTOM_DIAMETER_CM[phantom] / 2.0
    iso = np.asarray(rot["isocenter"], dtype=float)
    ax = _AXIS[rot.get("axis", "z")]
    lo, hi = iso - radius - 2.0, iso + radius + 2.0
    lo[ax] = iso[ax] - PHANTOM_LENGTH_CM / 2.0 - 2.0
    hi[ax] = iso[ax] + PHANTOM_LENGTH_CM / 2.0 + 2.0
    grid = VoxelGrid.from_bbox(lo, hi, GRID_RESOLUTION_CM)
    eng = engine or default_engine()
    eng.set_geometry(geometry)
    eng.bind_grid(grid.origin_cm, grid.shape, grid.voxel_size_cm, False)
    fl = N.RUN_FLUORESCENCE | N.RUN_GRID
    eng.run(SourceRecord(src), int(n_histories), _seed_from(seed), 0, 200_000, fl)
    grid.kerma_keV[...] = eng.read_grid(grid.shape, moments=False)["kerma"]
    dose = dose_map_Gy(grid, geometry) / float(n_histories)
    return ctdi_from_dose(dose, grid, iso, ax, radius)


def ctdi_from_dose(dose: np.ndarray, grid: VoxelGrid, iso, ax: int, radius: float):
    """(CTDIw, D_center, D_periphery) from a dose map: mean dose of the voxels within 2 cm of the
    axis and of those 1-2 cm below the surface, both within half the chamber length of the
    rotation centre along the axis."""
    iso = np.asarray(iso, dtype=float)
    plane = [k for k in range(3) if k != ax]
    c = grid.voxel_centers().reshape(*grid.shape, 3)
    r = np.hypot(c[..., plane[0]] - iso[plane[0]], c[..., plane[1]] - iso[plane[1]])
    z_ok = np.abs(c[..., ax] - iso[ax]) <= CHAMBER_LENGTH_CM / 2.0
    d_center = float(dose[z_ok & (r < 2.0)].mean())
    d_periph = float(dose[z_ok & (r >= radius - 2.0) & (r <= radius - 1.0)].mean())
    return d_center / 3.0 + 2.0 * d_periph / 3.0, d_center, d_periph


def effective_histories_from_ctdi(src: dict, seed: int | None = None, engine=None) -> float:
    """Effective photon count: the console CTDIvol [mGy] over CTDIw per history."""
    ctdiw, _, _ = ctdi_per_history_Gy(src, phantom=src.get("ctdi_phantom", "body"), seed=seed, engine=engine)
    return float(src["ctdi_vol_mGy"]) * 1e-3 / ctdiw
