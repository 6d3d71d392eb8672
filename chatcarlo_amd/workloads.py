"""Benchmark and parity workloads as scene dictionaries (scene.yaml equivalents).

Each function returns (scene_raw, grid_spec) where grid_spec is
(origin_cm, shape, voxel_cm) or None for the reference's default grid over the
world box.  Spectra are materialised as explicit `source.spectrum` lists
(Kramers + Al on fixed bins) so every consumer reads identical bins.
"""
from __future__ import annotations

import copy

import numpy as np

from .spectrum import kramers_spectrum_bins


def water_slab_60kev():
    """60 keV pencil beam (10 um square parallel beam on the axis) along +z into a 30x30x10 cm
    water slab in a 100 cm air cube; fluorescence on; 1 cm kerma grid over the slab.

    The beam axis runs along the corner line of four voxels; its 10 um width is resolvable in
    float32 at the grid's 15 cm offset, so GPU and FP64 oracle split photons alike."""
    raw = {
        "source": {"position": [0.0, 0.0, -15.0], "direction": [0.0, 0.0, 1.0],
                   "field": {"shape": "parallel", "size_cm": [1e-3, 1e-3]},
                   "spectrum": [{"energy_keV": 60.0, "weight": 1.0}]},
        "geometry": [
            {"name": "room_air", "shape": "box", "material": "air", "center": [0, 0, 0], "size_cm": [100, 100, 100]},
            {"name": "slab", "shape": "box", "material": "water", "center": [0, 0, 0], "size_cm": [30, 30, 10]},
        ],
        "physics": {"fluorescence": True},
    }
    return raw, (np.array([-15.0, -15.0, -5.0]), (30, 30, 10), 1.0)


def chest_room(kvp=80.0, filtration_mm_al=2.5):
    """400x400x300 cm air room with 20 cm concrete walls, a 30x20x50 cm water patient holding a
    3 cm x 40 cm bone cylinder, a 0.05 cm Pb apron sheet 100 cm lateral of the beam axis,
    80 kVp Kramers + 2.5 mm Al on 1 keV bins, 35x43 cm field at SID 180 cm; kerma + H*(10) on
    5 cm voxels over the whole world box (the reference's dose-grid semantics)."""
    raw = {
        "source": {"position": [0.0, -150.0, 130.0], "direction": [0.0, 1.0, 0.0],
                   "field": {"shape": "rect", "size_cm": [35.0, 43.0], "sid_cm": 180.0},
                   "spectrum": kramers_spectrum_bins(kvp, filtration_mm_al, 1.0)},
        "geometry": [
            {"name": "walls", "shape": "box", "material": "concrete", "center": [0, 0, 150],
             "size_cm": [440, 440, 340]},
            {"name": "room", "shape": "box", "material": "air", "center": [0, 0, 150], "size_cm": [400, 400, 300]},
            {"name": "patient", "shape": "box", "material": "water", "center": [0, 0, 130], "size_cm": [30, 20, 50]},
            {"name": "spine", "shape": "cylinder", "material": "bone", "center": [0, 5, 130], "radius_cm": 1.5,
             "height_cm": 40, "axis": "z"},
            {"name": "apron", "shape": "box", "material": "lead", "center": [100, 0, 120],
             "size_cm": [0.05, 60, 100]},
        ],
        "physics": {"fluorescence": True},
    }
    return raw, None


def ct_helical(kvp=120.0, filtration_mm_al=6.0):
    """32 cm x 15 cm PMMA CTDI phantom with five 1.3 cm holes in air; 120 kVp Kramers + 6 mm Al;
    continuous rotation at 60 cm radius; 4 cm fan collimation; 10 cm helical scan; 0.5 cm kerma grid
    over the phantom +- 2 cm."""
    geoms = [{"name": "phantom", "shape": "cylinder", "material": "pmma", "center": [0, 0, 0],
              "radius_cm": 16.0, "height_cm": 15.0, "axis": "z"}]
    for i, (x, y) in enumerate([(0, 0), (15, 0), (-15, 0), (0, 15), (0, -15)]):   # holes 1 cm below the surface
        geoms.append({"name": f"hole{i}", "shape": "cylinder", "material": "air", "center": [x, y, 0],
                      "radius_cm": 0.65, "height_cm": 15.0, "axis": "z"})
    raw = {
        "source": {"position": [0.0, -60.0, 0.0], "direction": [0.0, 1.0, 0.0],
                   "field": {"shape": "rect", "size_cm": [50.0, 4.0], "sid_cm": 60.0},
                   "rotation": {"isocenter": [0.0, 0.0, 0.0], "axis": "z", "scan_length_cm": 10.0},
                   "spectrum": kramers_spectrum_bins(kvp, filtration_mm_al, 1.0)},
        "geometry": geoms,
        "physics": {"fluorescence": True},
    }
    return raw, (np.array([-18.0, -18.0, -9.5]), (72, 72, 38), 0.5)


def angio_room(kvp=90.0, filtration_mm_al=3.0):
    """The chest room with a 30x20x180 cm water patient on a 1 cm PMMA table, an under-table
    90 kVp source (Kramers + 3 mm Al), and a 0.05 cm Pb ceiling shield above the operator
    position; the operator tally region is a 20 cm air sphere 60 cm lateral of the table."""
    raw = {
        "source": {"position": [0.0, 0.0, 20.0], "direction": [0.0, 0.0, 1.0],
                   "field": {"shape": "rect", "size_cm": [20.0, 20.0], "sid_cm": 100.0},
                   "spectrum": kramers_spectrum_bins(kvp, filtration_mm_al, 1.0)},
        "geometry": [
            {"name": "walls", "shape": "box", "material": "concrete", "center": [0, 0, 150],
             "size_cm": [440, 440, 340]},
            {"name": "room", "shape": "box", "material": "air", "center": [0, 0, 150], "size_cm": [400, 400, 300]},
            {"name": "table", "shape": "box", "material": "pmma", "center": [0, 0, 89.5], "size_cm": [50, 200, 1]},
            {"name": "patient", "shape": "box", "material": "water", "center": [0, 0, 100], "size_cm": [30, 180, 20]},
            {"name": "shield", "shape": "box", "material": "lead", "center": [45, 0, 125],
             "size_cm": [30, 40, 0.05]},
            {"name": "operator", "shape": "sphere", "material": "air", "center": [60, 0, 150], "radius_cm": 10.0},
        ],
        "physics": {"fluorescence": True},
    }
    return raw, None


def angio_region():
    """Operator H*(10) scoring region of angio_room: a 20 cm air sphere at 60 cm lateral, 150 cm high."""
    return np.array([60.0, 0.0, 150.0]), 10.0


def batched_scenes(n_scenes=64, master_seed=2026, max_prims=40):
    """64 random scenes: 1-40 box/cylinder/sphere primitives (water, bone, lead, concrete, pmma;
    sizes 0.05-50 cm) in a 300 cm air room; kVp U(50, 150) Kramers + 2.5 mm Al; histories
    log-uniform in 1e5-1e7; 2 cm grid over the room."""
    out = []
    mats = ["water", "bone", "lead", "concrete", "pmma"]
    for ss in np.random.SeedSequence(master_seed).spawn(n_scenes):
        rng = np.random.default_rng(ss)
        geoms = [{"name": "room", "shape": "box", "material": "air", "center": [0, 0, 0],
                  "size_cm": [300, 300, 300]}]
        for i in range(int(rng.integers(1, max_prims + 1))):
            shape = ["box", "cylinder", "sphere"][int(rng.integers(0, 3))]
            size = np.exp(rng.uniform(np.log(0.05), np.log(50.0), 3))
            c = rng.uniform(-100, 100, 3).tolist()
            g = {"name": f"p{i}", "shape": shape, "material": mats[int(rng.integers(0, 5))], "center": c}
            if shape == "box":
                g["size_cm"] = size.tolist()
            elif shape == "sphere":
                g["radius_cm"] = float(size[0] / 2)
            else:
                g.update(radius_cm=float(size[0] / 2), height_cm=float(size[1]), axis="xyz"[int(rng.integers(0, 3))])
            geoms.append(g)
        kvp = float(rng.uniform(50, 150))
        n_hist = int(np.exp(rng.uniform(np.log(1e5), np.log(1e7))))
        raw = {"source": {"position": [0.0, -140.0, 0.0], "direction": [0.0, 1.0, 0.0],
                          "field": {"shape": "rect", "size_cm": [30.0, 30.0], "sid_cm": 140.0},
                          "spectrum": kramers_spectrum_bins(kvp, 2.5, 1.0)},
               "geometry": geoms, "physics": {"fluorescence": True}}
        out.append((raw, (np.array([-150.0, -150.0, -150.0]), (150, 150, 150), 2.0), n_hist,
                    int(rng.integers(0, 2 ** 62))))
    return out


def copy_raw(raw):
    return copy.deepcopy(raw)


def _plain(x):
    """numpy scalars and arrays to plain Python values (for YAML)."""
    if isinstance(x, dict):
        return {k: _plain(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [_plain(v) for v in x]
    if isinstance(x, np.ndarray):
        return [_plain(v) for v in x.tolist()]
    if isinstance(x, np.generic):
        return x.item()
    return x


def _grid_doc(gs):
    if gs is None:
        return {"over": "world box", "resolution_cm": 5.0}
    o, s, v = gs
    return {"origin_cm": _plain(o), "shape": _plain(s), "voxel_cm": float(v)}


def fixture_documents():
    """The workloads as YAML documents: {file name: document}.

    Each single scene is a plain scene.yaml (source, geometry, physics) that the CLI and
    `load_scene` read. `workloads.yaml` indexes them with each workload's histories, seed and grid.
    `batched_64.yaml` holds the 64 batched scenes, each with its grid, histories and seed."""
    docs = {}
    index = {}
    singles = [("water_slab_60kev", water_slab_60kev(), 100_000, 42),
               ("chest_room", chest_room(), 10_000_000, 42),
               ("ct_helical", ct_helical(), 100_000_000, 42),
               ("angio_room", angio_room(), None, None)]
    for name, (raw, gs), n, seed in singles:
        docs[f"{name}.yaml"] = _plain(raw)
        index[name] = {"scene": f"{name}.yaml", "grid": _grid_doc(gs), "n_histories": n, "seed": seed}
    c, r = angio_region()
    index["angio_room"].update(grid=None, region={"center_cm": _plain(c), "radius_cm": float(r)},
                               target_rel_err=0.02)
    docs["batched_64.yaml"] = {"scenes": [
        {"scene": _plain(raw), "grid": _grid_doc(gs), "n_histories": int(n), "seed": int(s)}
        for raw, gs, n, s in batched_scenes(64)]}
    index["batched_64"] = {"scenes": "batched_64.yaml", "n_scenes": 64, "master_seed": 2026}
    docs["workloads.yaml"] = index
    return docs


def write_fixtures(directory):
    """Writes `fixture_documents()` as YAML files into `directory`."""
    import os

    import yaml
    os.makedirs(directory, exist_ok=True)
    for name, doc in fixture_documents().items():
        with open(os.path.join(directory, name), "w") as f:
            yaml.safe_dump(doc, f, sort_keys=False, default_flow_style=None, width=110)


if __name__ == "__main__":
    import sys
    write_fixtures(sys.argv[1] if len(sys.argv) > 1 else "fixtures/scenes")
