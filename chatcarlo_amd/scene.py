"""scene.yaml loading and validation (drop-in for chatcarlo/scene.py:18-309).

Errors are returned as data with a path ("geometry[2].size_cm"), never raised,
so a caller can fix a scene in a loop.  Units cm, z up.  Normalises
source.direction, field.shape, anode_angle_deg, rotation.axis and the
cylinder axis, and enforces the 150 keV upper limit of the tables.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import yaml

VALID_SHAPES = {"box", "cylinder", "sphere"}
VALID_AXES = {"x", "y", "z"}
E_MAX_KEV = 150.0


@dataclass
class SceneError:
    path: str
    message: str

    def __str__(self) -> str:
        return f"{self.path}: {self.message}"


@dataclass
class Scene:
    raw: dict
    errors: list = field(default_factory=list)
    warnings: list = field(default_factory=list)

    @property
    def ok(self) -> bool:
        return not self.errors


def _num(x) -> bool:
    return isinstance(x, (int, float)) and not isinstance(x, bool)


def _vec3(v, path, errors, name="vector"):
    if not (isinstance(v, (list, tuple)) and len(v) == 3 and all(_num(x) for x in v)):
        errors.append(SceneError(path, f"{name} must be a list of 3 numbers (e.g. [0, 0, 100])"))
        return None
    return [float(x) for x in v]


def load_scene(path: str) -> Scene:
    with open(path, encoding="utf-8") as f:
        raw = yaml.safe_load(f)
    return validate_scene(raw)


def _validate_source(src, errors, warnings):
    kvp = src.get("kvp")
    spectrum = src.get("spectrum")
    if spectrum is not None:
        if not (isinstance(spectrum, list) and spectrum and all(
                isinstance(s, dict) and _num(s.get("energy_keV")) and s["energy_keV"] > 0
                and _num(s.get("weight")) and s["weight"] > 0 for s in spectrum)):
            errors.append(SceneError("source.spectrum",
                                     "spectrum must be [{energy_keV: >0, weight: >0}, ...] "
                                     "(one entry for a monoenergetic beam)"))
        elif any(s["energy_keV"] > E_MAX_KEV for s in spectrum):
            bad = max(s["energy_keV"] for s in spectrum)
            errors.append(SceneError("source.spectrum",
                                     f"energy_keV={bad!r} exceeds the {E_MAX_KEV:g} keV diagnostic limit"))
        if kvp is not None:
            errors.append(SceneError("source", "source.spectrum and source.kvp cannot be combined"))
        for key in ("mas", "ctdi_vol_mGy"):
            if src.get(key) is not None:
                errors.append(SceneError("source", f"source.spectrum and source.{key} cannot be combined"))
        if src.get("heel_effect"):
            errors.append(SceneError("source", "source.spectrum and source.heel_effect cannot be combined"))
    elif not _num(kvp) or not (20 <= kvp <= 150):
        errors.append(SceneError("source.kvp", f"kvp={kvp!r}: give a number in 20-150 kV "
                                               "(or give source.spectrum)"))
    _vec3(src.get("position"), "source.position", errors, "focal spot position")
    dirv = _vec3(src.get("direction"), "source.direction", errors, "central axis direction")
    if dirv is not None:
        n = math.sqrt(sum(x * x for x in dirv))
        if n < 1e-9:
            errors.append(SceneError("source.direction", "direction vector is zero"))
        else:
            src["direction"] = [x / n for x in dirv]
    fld = src.get("field")
    if fld is None:
        errors.append(SceneError("source.field", "field is missing"))
    else:
        fshape = fld.get("shape", "rect")
        fld["shape"] = fshape
        if fshape == "cone":
            dia = fld.get("diameter_cm")
            if not _num(dia) or dia <= 0:
                errors.append(SceneError("source.field.diameter_cm",
                                         "a cone field needs a positive diameter_cm at the SID plane"))
        elif fshape in ("rect", "parallel"):
            size = fld.get("size_cm")
            if not (isinstance(size, (list, tuple)) and len(size) == 2 and all(_num(x) and x > 0 for x in size)):
                errors.append(SceneError("source.field.size_cm", "field size must be [width, height] > 0"))
        else:
            errors.append(SceneError("source.field.shape",
                                     f"shape={fshape!r}: one of rect (default), cone, parallel"))
        if fshape == "parallel":
            for key in ("mas", "heel_effect"):
                if src.get(key):
                    errors.append(SceneError("source", f"a parallel field cannot be combined with source.{key}"))
        else:
            sid = fld.get("sid_cm")
            if not _num(sid) or sid <= 0:
                errors.append(SceneError("source.field.sid_cm", "SID must be a positive number"))
    filt = src.get("filtration_mm_al", 2.5)
    if not _num(filt) or filt < 0:
        errors.append(SceneError("source.filtration_mm_al", "total filtration must be >= 0 mm Al"))
    elif filt < 1.5 and _num(kvp) and kvp >= 70:
        warnings.append(SceneError("source.filtration_mm_al",
                                   f"total filtration {filt} mm Al is below the usual 2.5 mm Al requirement"))
    angle = src.get("anode_angle_deg", 12.0)
    if not _num(angle) or not (5.0 <= angle <= 20.0):
        warnings.append(SceneError("source.anode_angle_deg",
                                   f"anode angle {angle!r}: diagnostic tubes use 5-20 degrees (using 12)"))
        angle = 12.0
    src["anode_angle_deg"] = float(angle)
    mas = src.get("mas")
    if mas is not None:
        if not _num(mas) or mas <= 0:
            errors.append(SceneError("source.mas", "mas must be a positive number"))
        else:
            src["mas"] = float(mas)
    ctdi = src.get("ctdi_vol_mGy")
    if ctdi is not None:
        if not _num(ctdi) or ctdi <= 0:
            errors.append(SceneError("source.ctdi_vol_mGy", "ctdi_vol_mGy must be a positive number"))
        else:
            src["ctdi_vol_mGy"] = float(ctdi)
        if mas is not None:
            errors.append(SceneError("source", "mas and ctdi_vol_mGy cannot be combined"))
        if src.get("rotation") is None:
            errors.append(SceneError("source.ctdi_vol_mGy", "CTDIvol calibration needs source.rotation"))
        ph = src.get("ctdi_phantom", "body")
        src["ctdi_phantom"] = ph
        if ph not in ("body", "head"):
            errors.append(SceneError("source.ctdi_phantom", f"ctdi_phantom={ph!r}: body or head"))
    if src.get("heel_effect"):
        ad = _vec3(src.get("anode_direction"), "source.anode_direction", errors,
                   "anode direction (required with heel_effect)")
        if ad is not None and isinstance(src.get("direction"), list):
            dv = src["direction"]
            dot = sum(ad[k] * dv[k] for k in range(3))
            perp = [ad[k] - dot * dv[k] for k in range(3)]
            if math.sqrt(sum(x * x for x in perp)) < 1e-6:
                errors.append(SceneError("source.anode_direction",
                                         "anode direction is parallel to the central axis"))
    rot = src.get("rotation")
    if rot is not None:
        if not isinstance(rot, dict):
            errors.append(SceneError("source.rotation", "rotation must be a mapping"))
        else:
            _vec3(rot.get("isocenter"), "source.rotation.isocenter", errors, "isocenter")
            axis = rot.get("axis", "z")
            rot["axis"] = axis
            if axis not in VALID_AXES:
                errors.append(SceneError("source.rotation.axis", f"axis={axis!r}: one of x/y/z"))
            n_angles = rot.get("n_angles")
            if n_angles is not None and not (isinstance(n_angles, int) and n_angles >= 2):
                errors.append(SceneError("source.rotation.n_angles", "n_angles must be an integer >= 2"))
            scan = rot.get("scan_length_cm")
            if scan is not None:
                if not _num(scan) or scan <= 0:
                    errors.append(SceneError("source.rotation.scan_length_cm",
                                             "scan_length_cm must be a positive number"))
                else:
                    rot["scan_length_cm"] = float(scan)


def validate_scene(raw: dict) -> Scene:
    """chatcarlo/scene.py:52-309."""
    errors: list = []
    warnings: list = []
    scene = Scene(raw=raw, errors=errors, warnings=warnings)
    if not isinstance(raw, dict):
        errors.append(SceneError("(root)", "the YAML top level must be a mapping"))
        return scene
    src = raw.get("source")
    if src is None:
        errors.append(SceneError("source", "source section is missing"))
    else:
        _validate_source(src, errors, warnings)
    geoms = raw.get("geometry")
    if not isinstance(geoms, list) or not geoms:
        errors.append(SceneError("geometry", "geometry must list at least one object"))
        geoms = []
    names = set()
    for i, g in enumerate(geoms):
        p = f"geometry[{i}]"
        if not isinstance(g, dict):
            errors.append(SceneError(p, "each object must be a mapping"))
            continue
        name = g.get("name", f"object_{i}")
        g["name"] = name
        if name in names:
            errors.append(SceneError(f"{p}.name", f"object name '{name}' is duplicated"))
        names.add(name)
        shape = g.get("shape")
        if shape not in VALID_SHAPES:
            errors.append(SceneError(f"{p}.shape", f"shape={shape!r}: one of {sorted(VALID_SHAPES)}"))
            continue
        if not g.get("material"):
            errors.append(SceneError(f"{p}.material", "material is missing"))
        _vec3(g.get("center"), f"{p}.center", errors, "center")
        if shape == "box":
            size = g.get("size_cm")
            if not (isinstance(size, (list, tuple)) and len(size) == 3 and all(_num(x) and x > 0 for x in size)):
                errors.append(SceneError(f"{p}.size_cm", "a box needs size_cm [x, y, z] > 0"))
        elif shape == "cylinder":
            if not (_num(g.get("radius_cm")) and g["radius_cm"] > 0):
                errors.append(SceneError(f"{p}.radius_cm", "a cylinder needs radius_cm > 0"))
            if not (_num(g.get("height_cm")) and g["height_cm"] > 0):
                errors.append(SceneError(f"{p}.height_cm", "a cylinder needs height_cm > 0"))
            axis = g.get("axis", "z")
            g["axis"] = axis
            if axis not in VALID_AXES:
                errors.append(SceneError(f"{p}.axis", f"axis={axis!r}: one of x/y/z"))
        elif shape == "sphere":
            if not (_num(g.get("radius_cm")) and g["radius_cm"] > 0):
                errors.append(SceneError(f"{p}.radius_cm", "a sphere needs radius_cm > 0"))
    physics = raw.get("physics")
    if physics is not None:
        if not isinstance(physics, dict):
            errors.append(SceneError("physics", "physics must be a mapping"))
        else:
            unknown = set(physics) - {"fluorescence"}
            if unknown:
                errors.append(SceneError("physics", f"unknown keys: {sorted(unknown)}"))
            fluor = physics.get("fluorescence", True)
            if not isinstance(fluor, bool):
                errors.append(SceneError("physics.fluorescence", "fluorescence must be true/false"))
            else:
                physics["fluorescence"] = fluor
    if scene.ok and src:
        pos = src["position"]
        for g in geoms:
            if g.get("shape") == "box" and "size_cm" in g and "center" in g:
                c, s = g["center"], g["size_cm"]
                if all(abs(pos[k] - c[k]) < s[k] / 2 for k in range(3)):
                    warnings.append(SceneError("source.position",
                                               f"the focal spot is inside object '{g['name']}'"))
    return scene
