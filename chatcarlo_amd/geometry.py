"""Scene geometry: the reference's Geometry API over a packed primitive list.

chatcarlo/geometry.py:121-231 keeps the scene's primitive list (box, sphere,
cylinder along x/y/z; later entries win where they overlap; background air),
puts the world box at the primitives' bounding box +- 50 cm and answers
material_at / next_boundary on the CPU.  Here the list is packed into the
fixed 48-byte records the GPU kernels read from LDS (cc_prim in
include/chatcarlo_hip.h); material_at / next_boundary run on the GPU through
the C-ABI.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

_AXES = {"x": 0, "y": 1, "z": 2}
KIND = {"box": 0, "sphere": 1, "cylinder": 2}
MAX_PRIMS = 64

PRIM_DTYPE = np.dtype([("kind", "<i4"), ("axis", "<i4"), ("material", "<i4"), ("pad", "<i4"),
                       ("p", "<f4", (8,))])
WORLD_DTYPE = np.dtype([("lo", "<f4", (3,)), ("hi", "<f4", (3,)), ("background", "<i4"),
                        ("n_prim", "<i4")])


def shape_bbox(g: dict):
    """Axis-aligned bounds of one primitive (chatcarlo/geometry.py:137-158)."""
    c = np.asarray(g["center"], dtype=float)
    shape = g["shape"]
    if shape == "box":
        half = np.asarray(g["size_cm"], dtype=float) / 2
        return c - half, c + half
    if shape == "sphere":
        r = float(g["radius_cm"])
        return c - r, c + r
    if shape == "cylinder":
        ai = _AXES[g.get("axis", "z")]
        r, h = float(g["radius_cm"]), float(g["height_cm"])
        lo, hi = c.copy(), c.copy()
        for k in range(3):
            ext = h / 2 if k == ai else r
            lo[k] -= ext
            hi[k] += ext
        return lo, hi
    raise ValueError(f"unknown shape: {shape}")


def pack_primitives(geoms, mat_id):
    """Scene geometry list -> PRIM_DTYPE records (float32 parameters)."""
    if len(geoms) > MAX_PRIMS:
        raise ValueError(f"at most {MAX_PRIMS} primitives are supported, got {len(geoms)}")
    arr = np.zeros(len(geoms), dtype=PRIM_DTYPE)
    for i, g in enumerate(geoms):
        shape = g["shape"]
        if shape not in KIND:
            raise ValueError(f"unknown shape: {shape}")
        arr[i]["kind"] = KIND[shape]
        arr[i]["material"] = mat_id(g["material"])
        c = np.asarray(g["center"], dtype=float)
        if shape == "box":
            lo, hi = shape_bbox(g)
            arr[i]["p"][0:3] = lo
            arr[i]["p"][3:6] = hi
        elif shape == "sphere":
            arr[i]["p"][0:3] = c
            arr[i]["p"][3] = float(g["radius_cm"])
        else:
            arr[i]["axis"] = _AXES[g.get("axis", "z")]
            arr[i]["p"][0:3] = c
            arr[i]["p"][3] = float(g["radius_cm"])
            arr[i]["p"][4] = float(g["height_cm"]) / 2.0
    return arr


@dataclass
class Geometry:
    """Drop-in for chatcarlo.geometry.Geometry (chatcarlo/geometry.py:121-135).

    geoms: the scene's geometry list; background: material outside every
    primitive; bbox_margin_cm: world box margin around the primitives.
    """
    geoms: list
    background: str = "air"
    bbox_margin_cm: float = 50.0
    bbox_min: np.ndarray = field(init=False)
    bbox_max: np.ndarray = field(init=False)

    def __post_init__(self):
        los, his = zip(*(shape_bbox(g) for g in self.geoms))
        self.bbox_min = np.min(np.stack(los), axis=0) - self.bbox_margin_cm
        self.bbox_max = np.max(np.stack(his), axis=0) + self.bbox_margin_cm

    def packed(self, mat_id):
        prims = pack_primitives(self.geoms, mat_id)
        world = np.zeros((), dtype=WORLD_DTYPE)
        world["lo"] = self.bbox_min
        world["hi"] = self.bbox_max
        world["background"] = mat_id(self.background)
        world["n_prim"] = len(prims)
        return prims, world

    def nearest_object_distance_cm(self, point):
        """chatcarlo/geometry.py:188-203 (host arithmetic, used for warnings only)."""
        if not self.geoms:
            return None
        pos = np.asarray(point, dtype=float)
        return min(float(np.linalg.norm(pos - np.clip(pos, *shape_bbox(g)))) for g in self.geoms)

    def material_at(self, points: np.ndarray) -> np.ndarray:
        """Points (N,3) -> material names (N,), evaluated on the GPU."""
        from .transport import default_engine
        eng = default_engine()
        ids = eng.locate(self, np.asarray(points, dtype=float))
        names = np.array(eng.pack.materials, dtype=object)
        return names[ids]

    def next_boundary(self, o: np.ndarray, d: np.ndarray, eps: float = 1e-6):
        """(t_boundary (N,), is_escape (N,)) evaluated on the GPU."""
        from .transport import default_engine
        t, esc, _ = default_engine().boundary(self, np.asarray(o, dtype=float),
                                              np.asarray(d, dtype=float), eps)
        return t.astype(float), esc.astype(bool)
