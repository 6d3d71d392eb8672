"""Write tests/golden/reference_runs.json from the reference's own committed run results.

The reference commits the outputs of its transport on simple scenes as data files.  This script
reads those data files (JSON and log lines, nothing is imported or run from the reference) and
stores, per scene, the scene parameters of the run script beside the values it produced:

* PDD/OCR in a water phantom, 60 keV parallel 10x10 cm beam, 47 box bins of kerma in Gy/history
  with their standard errors: docs/egs5_crosscheck/chatcarlo_pdd60_results.json, scene and bins
  of docs/egs5_crosscheck/run_chatcarlo_pdd60.py:37-100;
* K-fluorescence slabs (Pb 100 keV 0.05 cm, Pb 150 keV 0.135 cm, Cu 100 keV 0.75 cm, Cu 20 keV
  0.01 cm, W 100 keV 0.035 cm), pencil beam, fluorescence on and off:
  docs/egs5_crosscheck/fluorescence*/chatcarlo_results.json, scenes of the run_chatcarlo_fluorescence.py
  beside each (source gap and world margin differ between them);
* uncollided transmission of a water slab (20 keV 1.5 cm, 60 keV 10 cm, 150 keV 10 cm), pencil
  beam: the result lines of run_chatcarlo_water{20,60,150}.log, scenes of the scripts beside them.

Usage: python tests/golden/make_reference_runs.py [reference root]  (default /root/reference)
"""
from __future__ import annotations

import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

# (directory, material as the run script names it, energy keV, thickness cm, source gap cm,
#  world margin cm, peak band keV) -- run_chatcarlo_fluorescence.py of each directory
SLABS = [
    ("fluorescence", "lead", 100.0, 0.05, 5.0, 50.0, (72.0, 86.0)),
    ("fluorescence_lead_150kev", "lead", 150.0, 0.135, 0.001, 0.01, (72.0, 86.0)),
    ("fluorescence_copper", "copper", 100.0, 0.75, 5.0, 50.0, (7.8, 9.2)),
    ("fluorescence_copper_20kev", "copper", 20.0, 0.01, 0.001, 0.01, (7.8, 9.2)),
    ("fluorescence_tungsten", "W", 100.0, 0.035, 5.0, 50.0, (56.0, 70.0)),
]
# (log file, thickness cm, energy keV, source gap cm, world margin cm) -- run_chatcarlo_water*.py
WATER = [
    ("water20kev/run_chatcarlo_water20.log", 1.5, 20.0, 0.01, 0.01),
    ("run_chatcarlo_water60.log", 10.0, 60.0, 10.0, 50.0),
    ("water150kev/run_chatcarlo_water150.log", 10.0, 150.0, 0.01, 0.01),
]


def main(root: str) -> None:
    base = os.path.join(root, "docs", "egs5_crosscheck")
    out = {"source": "gakugaku3333/chatcarlo docs/egs5_crosscheck (reference runs, n and seed as stored)"}
    with open(os.path.join(base, "chatcarlo_pdd60_results.json")) as f:
        pdd = json.load(f)
    out["pdd60"] = {
        "scene": {"energy_keV": 60.0, "field_half_cm": 5.0, "source_z_cm": -1e-4,
                  "phantom_center_cm": [0.0, 0.0, 10.0], "phantom_size_cm": [30.0, 30.0, 20.0],
                  "material": "water", "world_margin_cm": 50.0, "n_histories": 5_000_000, "seed": 1},
        "bins": {k: {"lo_cm": v["lo_cm"], "hi_cm": v["hi_cm"], "mass_g": v["mass_g"],
                     "mean_Gy": v["mean_Gy_per_history"], "sem_Gy": v["sem_Gy_per_history"]}
                 for k, v in pdd.items()},
    }
    slabs = {}
    for d, mat, e, t, gap, margin, band in SLABS:
        with open(os.path.join(base, d, "chatcarlo_results.json")) as f:
            r = json.load(f)
        runs = {}
        for key in ("off", "on"):
            v = r[key]
            assert v["energy_keV"] == e and v["thickness_cm"] == t, (d, key)
            runs[key] = {k: v[k] for k in ("n_histories", "seed", "linear_mu_per_cm", "n_fluorescence",
                                          "n_escaped", "uncollided_transmission",
                                          "uncollided_transmission_stderr", "transmitted_energy_fraction",
                                          "transmitted_energy_fraction_stderr",
                                          "peak_band_fraction_of_escaped")}
        slabs[d] = {"material": mat, "energy_keV": e, "thickness_cm": t, "source_gap_cm": gap,
                    "world_margin_cm": margin, "slab_size_cm": 100.0, "peak_band_keV": list(band),
                    "runs": runs}
    out["fluorescence_slabs"] = slabs
    water = {}
    pat = re.compile(r"n=(\d+), seed=(\d+)\) = ([0-9.]+)%\s+\(.*?±([0-9.]+)pp\)")
    pat_mu = re.compile(r"mu\(xraylib, water, [0-9.]+ keV\) = ([0-9.]+) /cm")
    for log, t, e, gap, margin in WATER:
        with open(os.path.join(base, log), encoding="utf-8") as f:
            txt = f.read()
        m = pat.search(txt)
        mu = pat_mu.search(txt)
        assert m and mu, log
        water[f"water_{int(e)}kev"] = {
            "material": "water", "energy_keV": e, "thickness_cm": t, "source_gap_cm": gap,
            "world_margin_cm": margin, "slab_size_cm": 100.0, "n_histories": int(m.group(1)),
            "seed": int(m.group(2)), "linear_mu_per_cm": float(mu.group(1)),
            "uncollided_transmission": float(m.group(3)) / 100.0,
            "uncollided_transmission_stderr": float(m.group(4)) / 100.0}
    out["water_transmission"] = water
    with open(os.path.join(HERE, "reference_runs.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
