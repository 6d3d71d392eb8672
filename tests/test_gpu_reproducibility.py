"""Same seed, same result across processes (the reference's tests/test_reproducibility.py:40-54):
two processes with different string-hash seeds (PYTHONHASHSEED changes set and dict iteration of
strings) run the same source sampling and transport_photons call on the GPU.  Every random number
comes from Philox keyed by the seed and history, and materials have fixed pack indices, so the
counters agree exactly; the deposits are FP64 atomic sums, so they agree to summation order."""
import json
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent

_SCRIPT = """
import json, sys
sys.path.insert(0, sys.argv[1])
import numpy as np
from chatcarlo_amd.geometry import Geometry
from chatcarlo_amd.source import sample_source_photons
from chatcarlo_amd.transport import transport_photons

src = {"kvp": 80, "position": [0, -30, 0], "direction": [0, 1, 0], "field": {"size_cm": [10, 10], "sid_cm": 30}}
geometry = Geometry([
    {"name": "slab", "shape": "box", "material": "water", "center": [0, 0, 0], "size_cm": [10, 4, 10]},
    {"name": "plate", "shape": "box", "material": "aluminum", "center": [0, 5, 0], "size_cm": [10, 1, 10]},
])
rng = np.random.default_rng(7)
pos, dirv, e = sample_source_photons(src, 3000, rng)
r = transport_photons(pos, dirv, e, geometry, rng)
print(json.dumps({"edep": sorted(r.energy_deposited.items()),
                  "counts": [int(r.absorbed.sum()), int(r.escaped.sum()), int(r.n_scatter.sum())],
                  "final": float(np.sum(r.final_energy))}))
"""


def test_same_seed_reproducible_across_hash_seeds():
    outs = []
    for hash_seed in ("0", "12345"):
        env = {**os.environ, "PYTHONHASHSEED": hash_seed}
        p = subprocess.run([sys.executable, "-c", _SCRIPT, str(ROOT)], capture_output=True, text=True,
                           env=env, check=True, timeout=120)
        outs.append(json.loads(p.stdout.strip().splitlines()[-1]))
    a, b = outs
    assert a["counts"] == b["counts"] and a["counts"][0] + a["counts"][1] == 3000
    assert [k for k, _ in a["edep"]] == [k for k, _ in b["edep"]]
    assert {"aluminum", "water"} <= {k for k, _ in a["edep"]}
    for (_, x), (_, y) in zip(a["edep"], b["edep"]):
        assert y == pytest.approx(x, rel=1e-12)
    assert a["final"] == b["final"]
