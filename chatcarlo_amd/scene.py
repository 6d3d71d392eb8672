"""scene.yaml loading and validation (the drop-in for chatcarlo/scene.py:18-309).

The rules a scene must meet are data here: tables of field specifications (what a value must
be, the error path and message when it is not, the default written back), a table of source
keys that cannot be combined, and per-shape tables for the geometry.  One small interpreter
(`_apply`) walks them.  Errors and warnings are returned as data with a path such as
"geometry[2].size_cm", never raised, so a caller can correct a scene in a loop.  Units cm, z up.

The verdicts and error paths are the reference's (its interface to callers that fix scenes),
including its notion of a number (any int or float, bool included) and its presence tests:
kvp, mas and ctdi_vol_mGy count as given when not None, heel_effect when true.  Normalised on
the way: source.direction (unit length), field.shape, anode_angle_deg, mas, ctdi_vol_mGy,
ctdi_phantom, rotation.axis and scan_length_cm, object names, the cylinder axis and
physics.fluorescence.  Spectra above 150 keV, the top of the cross-section tables, are errors.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any, Callable

import yaml

VALID_SHAPES = {"box", "cylinder", "sphere"}
VALID_AXES = {"x", "y", "z"}
E_MAX_KEV = 150.0
KVP_RANGE = (20.0, 150.0)


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


def load_scene(path: str) -> Scene:
    with open(path, encoding="utf-8") as f:
        raw = yaml.safe_load(f)
    return validate_scene(raw)


# ---- value predicates --------------------------------------------------------------------
def number(x) -> bool:
    return isinstance(x, (int, float))


def positive(x) -> bool:
    return number(x) and x > 0


def numbers(n: int, each: Callable[[Any], bool] = number) -> Callable[[Any], bool]:
    return lambda v: isinstance(v, (list, tuple)) and len(v) == n and all(each(x) for x in v)


def within(lo: float, hi: float) -> Callable[[Any], bool]:
    return lambda x: number(x) and lo <= x <= hi


def one_of(options) -> Callable[[Any], bool]:
    return lambda x: x in options


def _spectrum_form(s) -> bool:
    return isinstance(s, list) and bool(s) and all(
        isinstance(b, dict) and positive(b.get("energy_keV")) and positive(b.get("weight")) for b in s)


def _given(key):
    return lambda sec: sec.get(key) is not None


def _true(key):
    return lambda sec: bool(sec.get(key))


ABSENT = object()


@dataclass(frozen=True)
class Spec:
    """One value of a section.

    check: what the value must satisfy; message: the complaint (str.format with value=...).
    optional: a value that is absent or None is skipped.  default: used when the key is absent
    (an explicit None stays None and is checked), and written back when `write` is set.  store:
    the normalised form written back when the value passes.  warn: the complaint is a warning,
    and `fallback` replaces the value."""
    key: str
    check: Callable[[Any], bool]
    message: str
    optional: bool = False
    default: Any = ABSENT
    write: bool = False
    store: Callable[[Any], Any] | None = None
    warn: bool = False
    fallback: Any = ABSENT


def _apply(section: dict, prefix: str, specs, errors, warnings):
    """Check every spec against section; returns the keys whose values passed."""
    passed = set()
    for s in specs:
        v = section.get(s.key) if s.default is ABSENT else section.get(s.key, s.default)
        if v is None and s.optional:
            continue
        if s.write:
            section[s.key] = v
        if s.check(v):
            passed.add(s.key)
            if s.store is not None:
                section[s.key] = s.store(v)
            continue
        (warnings if s.warn else errors).append(SceneError(f"{prefix}.{s.key}", s.message.format(value=v)))
        if s.fallback is not ABSENT:
            section[s.key] = s.store(s.fallback) if s.store else s.fallback
    return passed


# ---- source ------------------------------------------------------------------------------
SPECTRUM = [
    Spec("spectrum", _spectrum_form, "spectrum must be a non-empty list of {{energy_keV: > 0, weight: > 0}} "
         "entries (one entry for a monoenergetic beam, e.g. [{{energy_keV: 60, weight: 1.0}}])"),
]
KVP = [Spec("kvp", within(*KVP_RANGE), "kvp={value!r}: the tube voltage must be a number in 20-150 kV "
            "(or give source.spectrum instead)")]
BEAM = [
    Spec("position", numbers(3), "the focal spot must be 3 numbers, e.g. [0, 0, 100]"),
    Spec("direction", numbers(3), "the central axis must be 3 numbers, e.g. [0, 0, -1]"),
]
TUBE = [
    Spec("filtration_mm_al", lambda x: number(x) and x >= 0, "the total filtration must be a number >= 0 mm Al",
         default=2.5),
    Spec("anode_angle_deg", within(5.0, 20.0), "anode angle {value!r}: diagnostic tubes have 5-20 degrees; "
         "12 degrees is used", default=12.0, store=float, warn=True, fallback=12.0),
    Spec("mas", positive, "mas (tube current-time product) must be a positive number", optional=True,
         store=float),
    Spec("ctdi_vol_mGy", positive, "ctdi_vol_mGy (the console CTDIvol) must be a positive number",
         optional=True, store=float),
]
CTDI = [Spec("ctdi_phantom", one_of(("body", "head")), "ctdi_phantom={value!r}: body (32 cm) or head (16 cm)",
             default="body", write=True)]
FIELD_SHAPES = {
    "rect": [Spec("size_cm", numbers(2, positive), "the field size must be [width, height], both > 0"),
             Spec("sid_cm", positive, "the SID (focal spot to field plane) must be a positive number")],
    "cone": [Spec("diameter_cm", positive, "a cone field needs its opening diameter_cm (> 0) at the SID plane"),
             Spec("sid_cm", positive, "the SID (focal spot to field plane) must be a positive number")],
    "parallel": [Spec("size_cm", numbers(2, positive), "the field size must be [width, height], both > 0")],
}
ROTATION = [
    Spec("isocenter", numbers(3), "the isocentre must be 3 numbers"),
    Spec("axis", one_of(VALID_AXES), "axis={value!r}: one of x, y, z", default="z", write=True),
    Spec("n_angles", lambda x: isinstance(x, int) and x >= 2,
         "n_angles (discrete gantry angles) must be an integer >= 2; leave it out for a continuous rotation",
         optional=True),
    Spec("scan_length_cm", positive, "scan_length_cm (helical scan range) must be a positive number",
         optional=True, store=float),
]

# (applies when, path, message): combinations the source cannot take
CONFLICTS = [
    (lambda s: _given("spectrum")(s) and _given("kvp")(s), "source",
     "source.spectrum and source.kvp exclude each other: the spectrum replaces the kvp, give only one"),
    (lambda s: _given("spectrum")(s) and _given("mas")(s), "source",
     "source.spectrum cannot take source.mas: the mAs calibration follows the kvp spectrum, not the given one"),
    (lambda s: _given("spectrum")(s) and _given("ctdi_vol_mGy")(s), "source",
     "source.spectrum cannot take source.ctdi_vol_mGy"),
    (lambda s: _given("spectrum")(s) and _true("heel_effect")(s), "source",
     "source.spectrum cannot take source.heel_effect: the heel spectra follow the kvp"),
    (lambda s: _field_shape(s) == "parallel" and _given("mas")(s), "source",
     "a parallel field cannot take source.mas: the mAs calibration needs a focal spot and SID"),
    (lambda s: _field_shape(s) == "parallel" and _true("heel_effect")(s), "source",
     "a parallel field cannot take source.heel_effect"),
    (lambda s: _given("ctdi_vol_mGy")(s) and _given("mas")(s), "source",
     "source.mas and source.ctdi_vol_mGy exclude each other: mas for radiography, ctdi_vol_mGy for CT"),
    (lambda s: _given("ctdi_vol_mGy")(s) and not _given("rotation")(s), "source.ctdi_vol_mGy",
     "the CTDIvol calibration needs source.rotation (a rotating gantry)"),
]


def _field_shape(src):
    fld = src.get("field")
    return fld.get("shape", "rect") if isinstance(fld, dict) else None


def _spectrum_ceiling(src, errors):
    top = max(b["energy_keV"] for b in src["spectrum"])
    if top > E_MAX_KEV:
        errors.append(SceneError("source.spectrum", f"energy_keV={top!r} is above {E_MAX_KEV:g} keV, the top of "
                                                    "the cross-section tables"))


def _unit_direction(src, errors):
    d = [float(x) for x in src["direction"]]
    n = math.sqrt(sum(x * x for x in d))
    if n < 1e-9:
        errors.append(SceneError("source.direction", "the direction vector has zero length"))
    else:
        src["direction"] = [x / n for x in d]


def _anode_direction(src, errors):
    if not numbers(3)(src.get("anode_direction")):
        errors.append(SceneError("source.anode_direction", "heel_effect needs anode_direction, 3 numbers"))
        return
    d = src.get("direction")
    if not isinstance(d, list):
        return
    a = [float(x) for x in src["anode_direction"]]
    along = sum(a[k] * d[k] for k in range(3))
    if math.sqrt(sum((a[k] - along * d[k]) ** 2 for k in range(3))) < 1e-6:
        errors.append(SceneError("source.anode_direction",
                                 "the anode direction is parallel to the central axis; it needs a component "
                                 "across the beam"))


def _check_source(src, errors, warnings):
    if src.get("spectrum") is not None:
        if _apply(src, "source", SPECTRUM, errors, warnings):
            _spectrum_ceiling(src, errors)
    else:
        _apply(src, "source", KVP, errors, warnings)
    if "direction" in _apply(src, "source", BEAM, errors, warnings):
        _unit_direction(src, errors)
    fld = src.get("field")
    if fld is None:
        errors.append(SceneError("source.field", "the field is missing"))
    else:
        fld["shape"] = shape = fld.get("shape", "rect")
        specs = FIELD_SHAPES.get(shape)
        if specs is None:
            errors.append(SceneError("source.field.shape", f"shape={shape!r}: rect (default), cone or parallel"))
            specs = [FIELD_SHAPES["rect"][1]]   # an unknown shape is still a diverging field
        _apply(fld, "source.field", specs, errors, warnings)
    passed = _apply(src, "source", TUBE, errors, warnings)
    filt, kvp = src.get("filtration_mm_al", 2.5), src.get("kvp")
    if "filtration_mm_al" in passed and filt < 1.5 and number(kvp) and kvp >= 70:
        warnings.append(SceneError("source.filtration_mm_al", f"total filtration {filt} mm Al is below the "
                                                              "2.5 mm Al that diagnostic tubes usually need"))
    for applies, path, message in CONFLICTS:
        if applies(src):
            errors.append(SceneError(path, message))
    if src.get("ctdi_vol_mGy") is not None:
        _apply(src, "source", CTDI, errors, warnings)
    if src.get("heel_effect"):
        _anode_direction(src, errors)
    rot = src.get("rotation")
    if rot is not None:
        if isinstance(rot, dict):
            _apply(rot, "source.rotation", ROTATION, errors, warnings)
        else:
            errors.append(SceneError("source.rotation", "rotation must be a mapping"))


# ---- geometry ----------------------------------------------------------------------------
COMMON = [
    Spec("material", bool, "the material is missing"),
    Spec("center", numbers(3), "the centre must be 3 numbers"),
]
SHAPES = {
    "box": [Spec("size_cm", numbers(3, positive), "a box needs size_cm [x, y, z], all > 0")],
    "cylinder": [Spec("radius_cm", positive, "a cylinder needs radius_cm > 0"),
                 Spec("height_cm", positive, "a cylinder needs height_cm > 0"),
                 Spec("axis", one_of(VALID_AXES), "axis={value!r}: one of x, y, z", default="z", write=True)],
    "sphere": [Spec("radius_cm", positive, "a sphere needs radius_cm > 0")],
}


def _check_geometry(geoms, errors, warnings):
    seen = set()
    for i, g in enumerate(geoms):
        where = f"geometry[{i}]"
        if not isinstance(g, dict):
            errors.append(SceneError(where, "each object must be a mapping"))
            continue
        g["name"] = name = g.get("name", f"object_{i}")
        if name in seen:
            errors.append(SceneError(f"{where}.name", f"the object name '{name}' is used twice"))
        seen.add(name)
        specs = SHAPES.get(g.get("shape"))
        if specs is None:
            errors.append(SceneError(f"{where}.shape", f"shape={g.get('shape')!r}: one of {sorted(VALID_SHAPES)}"))
            continue
        _apply(g, where, COMMON + specs, errors, warnings)


# ---- physics -----------------------------------------------------------------------------
PHYSICS = [Spec("fluorescence", lambda x: isinstance(x, bool), "fluorescence must be true or false",
                default=True, store=bool)]


def _check_physics(physics, errors, warnings):
    if not isinstance(physics, dict):
        errors.append(SceneError("physics", "physics must be a mapping"))
        return
    unknown = sorted(set(physics) - {s.key for s in PHYSICS})
    if unknown:
        errors.append(SceneError("physics", f"unknown keys: {unknown}"))
    _apply(physics, "physics", PHYSICS, errors, warnings)


def _focal_spot_inside(src, geoms, warnings):
    p = src["position"]
    for g in geoms:
        if g.get("shape") == "box" and "size_cm" in g and "center" in g:
            if all(abs(p[k] - g["center"][k]) < g["size_cm"][k] / 2 for k in range(3)):
                warnings.append(SceneError("source.position", f"the focal spot is inside object '{g['name']}'"))


def validate_scene(raw: dict) -> Scene:
    """Checks and normalises a scene dict in place (chatcarlo/scene.py:52-309)."""
    scene = Scene(raw=raw)
    errors, warnings = scene.errors, scene.warnings
    if not isinstance(raw, dict):
        errors.append(SceneError("(root)", "the top level of the YAML must be a mapping"))
        return scene
    src = raw.get("source")
    if src is None:
        errors.append(SceneError("source", "the source section is missing"))
    else:
        _check_source(src, errors, warnings)
    geoms = raw.get("geometry")
    if not isinstance(geoms, list) or not geoms:
        errors.append(SceneError("geometry", "geometry must list at least one object"))
        geoms = []
    _check_geometry(geoms, errors, warnings)
    if raw.get("physics") is not None:
        _check_physics(raw["physics"], errors, warnings)
    if scene.ok and src:
        _focal_spot_inside(src, geoms, warnings)
    return scene
