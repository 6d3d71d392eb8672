"""CTDIvol calibration on the GPU: chatcarlo's own checks (tests/test_ctdi.py: finite and
positive, head phantom above body, periphery above centre in the body phantom, effective photon
count x CTDIw = CTDIvol), run_transport's n_photons_real, and the CLI's calibrated outputs."""
import functools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_SRC = {"kvp": 120.0, "filtration_mm_al": 2.5, "position": [0.0, 0.0, 170.0], "direction": [0.0, 0.0, -1.0],
        "rotation": {"isocenter": [0.0, 0.0, 110.0], "axis": "y", "scan_length_cm": 30.0},
        "field": {"shape": "rect", "size_cm": [50.0, 4.0], "sid_cm": 60.0}}


@functools.lru_cache(maxsize=None)
def _ctdi(phantom):
    from chatcarlo_amd.ctdi import ctdi_per_history_Gy
    return ctdi_per_history_Gy(_SRC, phantom=phantom, n_histories=400_000, seed=31)


def test_ctdi_positive_and_finite():
    for v in _ctdi("body"):
        assert np.isfinite(v) and v > 0


def test_head_above_body_and_periphery_above_centre():
    assert _ctdi("head")[0] > _ctdi("body")[0]
    _, dc, dp = _ctdi("body")
    assert dp > dc


def test_effective_histories_consistency():
    from chatcarlo_amd.ctdi import ctdi_per_history_Gy, effective_histories_from_ctdi
    src = {**_SRC, "ctdi_vol_mGy": 12.0, "ctdi_phantom": "body"}
    n_eff = effective_histories_from_ctdi(src, seed=31)
    ctdiw, _, _ = ctdi_per_history_Gy(src, phantom="body", n_histories=1_000_000, seed=77)
    assert n_eff * ctdiw * 1e3 == pytest.approx(12.0, rel=0.1)


def test_run_transport_calibrates_and_cli_writes_absolute_maps(tmp_path):
    import yaml
    from chatcarlo_amd.__main__ import main
    from chatcarlo_amd.scene import validate_scene
    from chatcarlo_amd.transport import run_transport
    raw = {"source": {**_SRC, "ctdi_vol_mGy": 10.0},
           "geometry": [{"name": "patient", "shape": "cylinder", "material": "water", "center": [0, 0, 110],
                         "radius_cm": 15.0, "height_cm": 60.0, "axis": "y"}]}
    sc = validate_scene(raw)
    assert sc.ok, sc.errors
    r = run_transport(sc, n_histories=100_000, seed=5, dose_grid=True, grid_resolution_cm=5.0)
    ctdiw, _, _ = _ctdi("body")
    assert r.n_photons_real == pytest.approx(10.0e-3 / ctdiw, rel=0.1)
    scene = tmp_path / "ct.yaml"
    scene.write_text(yaml.safe_dump(raw))
    out = tmp_path / "d.npz"
    assert main(["run", str(scene), "-n", "50000", "--seed", "3", "--dose-grid", "--dose-out", str(out)]) == 0
    z = np.load(out)
    assert "dose_Gy" in z and "h10_pSv" in z
    scale = z["dose_Gy"].max() / z["dose_per_history_Gy"].max()
    assert scale == pytest.approx(r.n_photons_real, rel=0.15)
